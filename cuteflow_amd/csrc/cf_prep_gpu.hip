// Mesh preprocessing on the GPU, with the results of the host code in
// cf_prep.cpp bit for bit:
//   - cf_build_boundary_device: the directed-edge neighbour match and the
//     boundary classes (the reference's tab_bound_tri + boundary_kind,
//     maillage.cuf:781-1110, which runs an all-pairs O(N^2) search):
//     3N edge keys (a << 32 | b) radix-sorted with their cell (a stable sort
//     of cell-ordered input, i.e. ordered by (key, cell) like the host's
//     sort), then one binary search per face for the reversed edge;
//   - prepare_device: the internal layout of cfprep::prepare (neighbour
//     codes and face matching, Hilbert cell order, tile adjacency, inflow
//     tiles, inflow/outflow cells) from sorts and per-cell kernels.
// Sorts, selections and reductions are hipCUB's; everything else is one
// thread per cell or per face.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <vector>

#include "cf_prep.h"

namespace {

typedef unsigned long long u64;

#define GCHK(x)                          \
    do {                                 \
        if ((x) != hipSuccess) return CF_ERR_HIP; \
    } while (0)

inline int blocks(size_t n, int b = 256) { return (int)((n + b - 1) / b); }

// Device buffers freed on scope exit.
struct DevBufs {
    std::vector<void *> p;
    template <typename T>
    bool alloc(T *&ptr, size_t n) {
        ptr = nullptr;
        if (hipMalloc((void **)&ptr, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return false;
        p.push_back(ptr);
        return true;
    }
    ~DevBufs() {
        for (void *q : p) (void)hipFree(q);
    }
};

// Private stream, synchronised and destroyed on scope exit.
struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
};

template <typename T>
bool up(T *d, const T *h, size_t n, hipStream_t s) {
    return n == 0 || hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s) == hipSuccess;
}
template <typename T>
bool down(T *h, const T *d, size_t n, hipStream_t s) {
    return n == 0 || hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, s) == hipSuccess;
}

// stable radix sort of (key, value) pairs over the given key bits
template <typename K, typename V>
int sort_pairs(const K *kin, K *kout, const V *vin, V *vout, int n, int end_bit, hipStream_t s, DevBufs &bufs) {
    size_t bytes = 0;
    GCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, n, 0, end_bit, s));
    char *tmp;
    if (!bufs.alloc(tmp, bytes)) return CF_ERR_NOMEM;
    GCHK(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout, n, 0, end_bit, s));
    return CF_OK;
}

// ---- boundary table ---------------------------------------------------------

__global__ void check_conn_kernel(const int *conn, size_t m, int nnt, int *err) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < m && (conn[i] < 1 || conn[i] > nnt)) *err = 1;
}

__global__ void edge_keys_kernel(const int *conn, int n, u64 *keys, int *vals) {
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= 3 * (size_t)n) return;
    const int i = (int)(e / 3), k = (int)(e % 3);
    const unsigned a = (unsigned)conn[3 * (size_t)i + k], b = (unsigned)conn[3 * (size_t)i + (k + 1) % 3];
    keys[e] = ((u64)a << 32) | b;
    vals[e] = i;
}

// Neighbour across each face: among the cells holding the reversed edge,
// the highest one other than the cell itself (the host keeps the last match
// in (key, cell) order); none: the cell itself (undefined face).
__global__ void match_kernel(const int *conn, int n, const u64 *skeys, const int *svals, int *boundary) {
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t m = 3 * (size_t)n;
    if (e >= m) return;
    const int i = (int)(e / 3), k = (int)(e % 3);
    const unsigned a = (unsigned)conn[3 * (size_t)i + k], b = (unsigned)conn[3 * (size_t)i + (k + 1) % 3];
    const u64 want = ((u64)b << 32) | a;
    size_t lo = 0, hi = m;
    while (lo < hi) {  // first key >= want
        const size_t mid = (lo + hi) >> 1;
        if (skeys[mid] < want) lo = mid + 1;
        else hi = mid;
    }
    size_t top = lo, hi2 = m;
    while (top < hi2) {  // first key > want
        const size_t mid = (top + hi2) >> 1;
        if (skeys[mid] <= want) top = mid + 1;
        else hi2 = mid;
    }
    int nb = i + 1;
    for (size_t j = top; j > lo; j--)
        if (svals[j - 1] != i) {
            nb = svals[j - 1] + 1;
            break;
        }
    boundary[(size_t)k * n + i] = nb;
    boundary[3 * (size_t)n + (size_t)k * n + i] = i + 1;
}

// first position of each node in a node list (pos starts at INT_MAX)
__global__ void first_pos_kernel(const int *list, int nd, int nnt, int *pos) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nd) return;
    const int v = list[q];
    if (v >= 1 && v <= nnt) atomicMin(&pos[v], q);
}

__global__ void fill_int_kernel(int *p, size_t n, int v) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// boundary classes of the faces left without a neighbour (cf_build_boundary)
__global__ void classify_kernel(const int *conn, int n, int ndi, int ndo, int ndw, const int *pi, const int *po,
                                const int *pw, const int *num_in, const int *num_out, int *boundary) {
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= 3 * (size_t)n) return;
    const int i = (int)(e / 3), k = (int)(e % 3);
    const size_t at = (size_t)k * n + i, at2 = 3 * (size_t)n + at;
    if (boundary[at] != i + 1) return;
    const int n1 = conn[3 * (size_t)i + k], n2 = conn[3 * (size_t)i + (k + 1) % 3];
    const bool i1 = pi[n1] != INT_MAX, i2 = pi[n2] != INT_MAX;
    const bool o1 = po[n1] != INT_MAX, o2 = po[n2] != INT_MAX;
    const bool w1 = pw[n1] != INT_MAX, w2 = pw[n2] != INT_MAX;
    if (ndi > 0 && n1 != n2 && i1 && i2) {
        const int q = min(pi[n1], pi[n2]);
        boundary[at] = -1;
        boundary[at2] = num_in ? num_in[q] : 1;
        return;
    }
    if (ndo > 0 && n1 != n2 && o1 && o2) {
        const int q = min(po[n1], po[n2]);
        boundary[at] = -2;
        boundary[at2] = num_out ? num_out[q] : 1;
        return;
    }
    if (ndw > 0 && n1 != n2 && w1 && w2) {
        boundary[at] = -3;
        return;
    }
    if (ndi > 0 && ndw > 0 && ((w1 && i2) || (i1 && w2))) boundary[at] = -3;
    if (ndo > 0 && ndw > 0 && ((w1 && o2) || (o1 && w2))) boundary[at] = -3;
}

// ---- internal layout ----------------------------------------------------------

struct LongEntree {
    double v[CF_MAX_IO];
};

// Caller-order neighbour codes with set_undefined_boudaries_as_walls applied,
// and the face of the neighbour that shares each edge (cfprep::prepare).
// Errors: the smallest (cell * 3 + face) that fails, with its status, as
// (index << 4 | code) in *err (code 1: CF_ERR_MESH, 2: CF_ERR_ARG), so that
// the status is the one the host's loop in cell order meets first.
__global__ void nb0_kernel(const int *conn, const int *b, int n, int inlet, LongEntree le, int4 *nb0, u64 *err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t N = (size_t)n;
    int c[3];
    for (int k = 0; k < 3; k++) {
        int v = b[k * N + i];
        const int id = b[3 * N + k * N + i];
        if (v == i + 1) v = -3;
        int code = 0;
        if (v > 0) {
            if (v > n) code = 1;
            c[k] = v - 1;
        } else if (v == -1 || v == -2) {
            if (id < 1 || id > CF_MAX_IO) code = 1;
            else if (v == -1 && inlet == 2 && !(le.v[id - 1] > 0.0)) code = 2;
            c[k] = -((v == -1 ? 1 : 2) | (id << 2));
        } else if (v == -3) {
            c[k] = -3;
        } else {
            code = 1;
        }
        if (code) {
            atomicMin(err, ((u64)(3 * (size_t)i + k) << 4) | (u64)code);
            return;
        }
    }
    int w = 0;
    for (int j = 0; j < 3; j++) {
        int tjv = j;
        const int cc = c[j];
        if (cc >= 0) {
            const int A = conn[3 * (size_t)i + j], B = conn[3 * (size_t)i + (j + 1) % 3];
            const int *q = conn + 3 * (size_t)cc;
            if ((q[0] == A && q[1] == B) || (q[1] == A && q[0] == B)) tjv = 0;
            else if ((q[1] == A && q[2] == B) || (q[2] == A && q[1] == B)) tjv = 1;
            else if ((q[2] == A && q[0] == B) || (q[0] == A && q[2] == B)) tjv = 2;
            else tjv = 0;
        }
        w |= tjv << (2 * j);
    }
    nb0[i] = make_int4(c[0], c[1], c[2], w);
}

#pragma clang fp contract(off)
// centroids as the host computes them: ((0 + xa) + xb) + xc, / 3
__global__ void centroid_kernel(const int *conn, const double *xy, int n, double *cx, double *cy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sx = 0.0, sy = 0.0;
    for (int k = 0; k < 3; k++) {
        const size_t v = (size_t)(conn[3 * (size_t)i + k] - 1);
        sx = sx + xy[3 * v];
        sy = sy + xy[3 * v + 1];
    }
    cx[i] = sx / 3.0;
    cy[i] = sy / 3.0;
}

__global__ void hilbert_key_kernel(const double *cx, const double *cy, int n, const double *ext, unsigned *key,
                                   int *idx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x0 = ext[0], x1 = ext[1], y0 = ext[2], y1 = ext[3];
    const double span = fmax(fmax(x1 - x0, y1 - y0), 1e-300);
    const double sc = 65535.0 / span;
    uint32_t x = (uint32_t)fmin(65535.0, fmax(0.0, (cx[i] - x0) * sc));
    uint32_t y = (uint32_t)fmin(65535.0, fmax(0.0, (cy[i] - y0) * sc));
    // Hilbert index of order 16 (the host's hilbert_d)
    const uint32_t n1 = 65535u;
    uint32_t d = 0;
    for (uint32_t s = 1u << 15; s > 0; s >>= 1) {
        const uint32_t rx = (x & s) ? 1u : 0u;
        const uint32_t ry = (y & s) ? 1u : 0u;
        d += s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = n1 - x;
                y = n1 - y;
            }
            const uint32_t t = x;
            x = y;
            y = t;
        }
    }
    key[i] = d;
    idx[i] = i;
}

__global__ void iota_kernel(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ void inverse_kernel(const int *perm, int n, int *inv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inv[perm[i]] = i;
}

// internal-order neighbour codes and connectivity; per caller cell, whether
// it has an inflow or outflow face; per tile, an inflow face
__global__ void permute_kernel(const int4 *nb0, const int *conn, const int *perm, const int *inv, int n, int tile,
                               int4 *nbr, int *conn_out, unsigned char *io_flag, unsigned char *always) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const int o = perm[m];
    int4 r = nb0[o];
    int c[3] = {r.x, r.y, r.z};
    bool io = false, in = false;
    for (int k = 0; k < 3; k++) {
        if (c[k] >= 0) c[k] = inv[c[k]];
        else if (((-c[k]) & 3) == 1) io = in = true;
        else if (((-c[k]) & 3) == 2) io = true;
    }
    nbr[m] = make_int4(c[0], c[1], c[2], r.w);
    for (int k = 0; k < 3; k++) conn_out[3 * (size_t)m + k] = conn[3 * (size_t)o + k];
    io_flag[o] = io ? 1 : 0;
    if (in) always[m / tile] = 1;
}

// (tile, neighbour tile) keys of the internal faces crossing a tile edge;
// the others get the largest key
__global__ void tile_pair_kernel(const int4 *nbr, int n, int tile, u64 *keys) {
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= 3 * (size_t)n) return;
    const int m = (int)(e / 3), k = (int)(e % 3);
    const int4 r = nbr[m];
    const int c = k == 0 ? r.x : (k == 1 ? r.y : r.z);
    u64 key = ~0ull;
    if (c >= 0) {
        const unsigned t = (unsigned)(m / tile), tt = (unsigned)(c / tile);
        if (tt != t) key = ((u64)t << 32) | tt;
    }
    keys[e] = key;
}

// CSR offsets: the first unique key of each tile
__global__ void tile_off_kernel(const u64 *ukeys, int nu, int ntiles, int *off, int *adj) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nu) adj[t] = (int)(unsigned)(ukeys[t] & 0xffffffffull);
    if (t > ntiles) return;
    const u64 want = (u64)t << 32;
    int lo = 0, hi = nu;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ukeys[mid] < want) lo = mid + 1;
        else hi = mid;
    }
    off[t] = lo;
}

__global__ void map_inv_kernel(const int *sel, int n, const int *inv, int *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = inv[sel[i]];
}

}  // namespace

extern "C" int cf_build_boundary_device(int nelt, int nnt, const int *conn, int ndi, const int *ndinput,
                                        const int *num_in, int ndo, const int *ndoutput, const int *num_out, int ndw,
                                        const int *ndwall, int *boundary) {
    if (nelt <= 0 || nnt <= 0 || !conn || !boundary || ndi < 0 || ndo < 0 || ndw < 0 || (ndi > 0 && !ndinput) ||
        (ndo > 0 && !ndoutput) || (ndw > 0 && !ndwall))
        return CF_ERR_ARG;
    if ((size_t)nelt > (size_t)INT32_MAX / 8) return CF_ERR_ARG;
    const size_t N = (size_t)nelt, M = 3 * N;
    DevBufs B;  // freed after the stream guard below has synchronised
    hipStream_t s;
    GCHK(hipStreamCreate(&s));
    StreamGuard guard{s};
    int *d_conn, *d_vals, *d_svals, *d_bnd, *d_pi, *d_po, *d_pw, *d_in, *d_out, *d_wall, *d_nin = nullptr,
                                                                                         *d_nout = nullptr, *d_err;
    u64 *d_keys, *d_skeys;
    if (!B.alloc(d_conn, M) || !B.alloc(d_vals, M) || !B.alloc(d_svals, M) || !B.alloc(d_keys, M) ||
        !B.alloc(d_skeys, M) || !B.alloc(d_bnd, 6 * N) || !B.alloc(d_pi, (size_t)nnt + 1) ||
        !B.alloc(d_po, (size_t)nnt + 1) || !B.alloc(d_pw, (size_t)nnt + 1) || !B.alloc(d_in, (size_t)ndi) ||
        !B.alloc(d_out, (size_t)ndo) || !B.alloc(d_wall, (size_t)ndw) || !B.alloc(d_err, 1) ||
        (num_in && !B.alloc(d_nin, (size_t)ndi)) || (num_out && !B.alloc(d_nout, (size_t)ndo)))
        return CF_ERR_NOMEM;
    const int zero = 0;
    if (!up(d_conn, conn, M, s) || !up(d_in, ndinput, (size_t)ndi, s) || !up(d_out, ndoutput, (size_t)ndo, s) ||
        !up(d_wall, ndwall, (size_t)ndw, s) || !up(d_err, &zero, 1, s) ||
        (num_in && !up(d_nin, num_in, (size_t)ndi, s)) || (num_out && !up(d_nout, num_out, (size_t)ndo, s)))
        return CF_ERR_HIP;
    hipLaunchKernelGGL(check_conn_kernel, dim3(blocks(M)), dim3(256), 0, s, d_conn, M, nnt, d_err);
    int err = 0;
    if (!down(&err, d_err, 1, s) || hipStreamSynchronize(s) != hipSuccess) return CF_ERR_HIP;
    if (err) return CF_ERR_MESH;
    hipLaunchKernelGGL(edge_keys_kernel, dim3(blocks(M)), dim3(256), 0, s, d_conn, nelt, d_keys, d_vals);
    int rc = sort_pairs(d_keys, d_skeys, d_vals, d_svals, (int)M, 64, s, B);
    if (rc) return rc;
    hipLaunchKernelGGL(match_kernel, dim3(blocks(M)), dim3(256), 0, s, d_conn, nelt, d_skeys, d_svals, d_bnd);
    int *lists[3] = {d_in, d_out, d_wall}, *pos[3] = {d_pi, d_po, d_pw}, nds[3] = {ndi, ndo, ndw};
    for (int l = 0; l < 3; l++) {
        hipLaunchKernelGGL(fill_int_kernel, dim3(blocks((size_t)nnt + 1)), dim3(256), 0, s, pos[l], (size_t)nnt + 1,
                           INT_MAX);
        if (nds[l] > 0)
            hipLaunchKernelGGL(first_pos_kernel, dim3(blocks((size_t)nds[l])), dim3(256), 0, s, lists[l], nds[l], nnt,
                               pos[l]);
    }
    hipLaunchKernelGGL(classify_kernel, dim3(blocks(M)), dim3(256), 0, s, d_conn, nelt, ndi, ndo, ndw, d_pi, d_po,
                       d_pw, (const int *)d_nin, (const int *)d_nout, d_bnd);
    GCHK(hipGetLastError());
    if (!down(boundary, d_bnd, 6 * N, s) || hipStreamSynchronize(s) != hipSuccess) return CF_ERR_HIP;
    return CF_OK;
}

namespace {

// ---- boundary table without device allocations ----------------------------
//
// cf_build_boundary_device_noalloc works in `scratch`, a fixed device array
// (each GPU the library runs on holds its own copy until the process ends):
// edge keys and cells twice over (the radix sort ping-pongs between them),
// the first positions of the nodes in the three node lists, an error flag
// and the sort's temporary storage.  Host code reaches it only through the
// address hipGetSymbolAddress returns for the current device, taken once per
// device on the first call.

constexpr size_t kNaMaxCells = size_t(1) << 17;
constexpr size_t kNaMaxNodes = size_t(1) << 17;
constexpr size_t kNaMaxEdges = 3 * kNaMaxCells;
constexpr size_t kNaSortTemp = size_t(4) << 20;

constexpr size_t na_round(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kNaKeys0 = 0;
constexpr size_t kNaKeys1 = kNaKeys0 + na_round(kNaMaxEdges * sizeof(u64));
constexpr size_t kNaVals0 = kNaKeys1 + na_round(kNaMaxEdges * sizeof(u64));
constexpr size_t kNaVals1 = kNaVals0 + na_round(kNaMaxEdges * sizeof(int));
constexpr size_t kNaPos = kNaVals1 + na_round(kNaMaxEdges * sizeof(int));
constexpr size_t kNaErr = kNaPos + na_round(3 * (kNaMaxNodes + 1) * sizeof(int));
constexpr size_t kNaTemp = kNaErr + 256;
constexpr size_t kNaBytes = kNaTemp + kNaSortTemp;

static __device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kNaMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[kNaMaxDevices];

// Device address of the current device's scratch array.
static int scratch_of_current_device(unsigned char **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return CF_ERR_HIP;
    if (dev < 0 || dev >= kNaMaxDevices) return CF_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "cuteflow_hip: hipGetSymbolAddress(scratch) on device %d failed: %s\n", dev,
                    hipGetErrorString(e));
            return CF_ERR_HIP;
        }
        g_scratch_ptr[dev] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[dev];
    return CF_OK;
}

}  // namespace

extern "C" int cf_build_boundary_device_noalloc(int nelt, int nnt, const int *conn, int ndi, const int *ndinput,
                                                const int *num_in, int ndo, const int *ndoutput, const int *num_out,
                                                int ndw, const int *ndwall, int *boundary, void *stream) {
    if (nelt <= 0 || nnt <= 0 || !conn || !boundary || ndi < 0 || ndo < 0 || ndw < 0 || (ndi > 0 && !ndinput) ||
        (ndo > 0 && !ndoutput) || (ndw > 0 && !ndwall))
        return CF_ERR_ARG;
    if ((size_t)nelt > kNaMaxCells || (size_t)nnt > kNaMaxNodes) {
        fprintf(stderr,
                "cuteflow_hip: cf_build_boundary_device_noalloc: %d cells, %d nodes exceed its fixed scratch "
                "(at most %zu cells and %zu nodes); use cf_build_boundary_device\n",
                nelt, nnt, kNaMaxCells, kNaMaxNodes);
        return CF_ERR_SIZE;
    }
    unsigned char *base = nullptr;
    int rc = scratch_of_current_device(&base);
    if (rc) return rc;
    const hipStream_t s = (hipStream_t)stream;
    const size_t N = (size_t)nelt, M = 3 * N, P = (size_t)nnt + 1;
    u64 *keys0 = (u64 *)(base + kNaKeys0), *keys1 = (u64 *)(base + kNaKeys1);
    int *vals0 = (int *)(base + kNaVals0), *vals1 = (int *)(base + kNaVals1);
    int *pos[3] = {(int *)(base + kNaPos), (int *)(base + kNaPos) + P, (int *)(base + kNaPos) + 2 * P};
    int *d_err = (int *)(base + kNaErr);
    void *temp = base + kNaTemp;
    // the scratch keeps the previous call's contents: clear the error flag
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, d_err, (size_t)1, 0);
    hipLaunchKernelGGL(check_conn_kernel, dim3(blocks(M)), dim3(256), 0, s, conn, M, nnt, d_err);
    int err = 0;
    GCHK(hipGetLastError());
    if (!down(&err, d_err, 1, s) || hipStreamSynchronize(s) != hipSuccess) return CF_ERR_HIP;
    if (err) return CF_ERR_MESH;
    hipLaunchKernelGGL(edge_keys_kernel, dim3(blocks(M)), dim3(256), 0, s, conn, nelt, keys0, vals0);
    hipcub::DoubleBuffer<u64> kbuf(keys0, keys1);
    hipcub::DoubleBuffer<int> vbuf(vals0, vals1);
    size_t bytes = 0;
    GCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kbuf, vbuf, (int)M, 0, 64, s));
    if (bytes > kNaSortTemp) {
        fprintf(stderr, "cuteflow_hip: cf_build_boundary_device_noalloc: the sort needs %zu bytes of temporary "
                        "storage, the scratch holds %zu\n", bytes, kNaSortTemp);
        return CF_ERR_SIZE;
    }
    GCHK(hipcub::DeviceRadixSort::SortPairs(temp, bytes, kbuf, vbuf, (int)M, 0, 64, s));
    hipLaunchKernelGGL(match_kernel, dim3(blocks(M)), dim3(256), 0, s, conn, nelt, (const u64 *)kbuf.Current(),
                       (const int *)vbuf.Current(), boundary);
    const int *lists[3] = {ndinput, ndoutput, ndwall};
    const int nds[3] = {ndi, ndo, ndw};
    for (int l = 0; l < 3; l++) {
        hipLaunchKernelGGL(fill_int_kernel, dim3(blocks(P)), dim3(256), 0, s, pos[l], P, INT_MAX);
        if (nds[l] > 0)
            hipLaunchKernelGGL(first_pos_kernel, dim3(blocks((size_t)nds[l])), dim3(256), 0, s, lists[l], nds[l], nnt,
                               pos[l]);
    }
    hipLaunchKernelGGL(classify_kernel, dim3(blocks(M)), dim3(256), 0, s, conn, nelt, ndi, ndo, ndw,
                       (const int *)pos[0], (const int *)pos[1], (const int *)pos[2], num_in, num_out, boundary);
    GCHK(hipGetLastError());
    GCHK(hipStreamSynchronize(s));
    return CF_OK;
}

namespace cfprep {

int prepare_device(const cf_mesh *mesh, const cf_params *params, int tile, bool keep_order, Prep &out) {
    const int n = mesh->nelt;
    const size_t N = (size_t)n;
    out.nelt = n;
    out.tile = tile;
    DevBufs B;  // freed after the stream guard below has synchronised
    hipStream_t s;
    GCHK(hipStreamCreate(&s));
    StreamGuard guard{s};
    int *d_conn, *d_bnd, *d_perm, *d_inv, *d_conn2, *d_idx, *d_sel, *d_nsel, *d_io, *d_adj, *d_off, *d_nu;
    int4 *d_nb0, *d_nbr;
    double *d_xy, *d_cx, *d_cy, *d_ext;
    unsigned *d_key, *d_skey;
    unsigned char *d_ioflag, *d_always;
    u64 *d_err, *d_tk, *d_stk, *d_uk;
    const int ntiles = (int)((N + tile - 1) / tile);
    if (!B.alloc(d_conn, 3 * N) || !B.alloc(d_bnd, 6 * N) || !B.alloc(d_perm, N) || !B.alloc(d_inv, N) ||
        !B.alloc(d_conn2, 3 * N) || !B.alloc(d_idx, N) || !B.alloc(d_sel, N) || !B.alloc(d_nsel, 1) ||
        !B.alloc(d_io, N) || !B.alloc(d_nb0, N) || !B.alloc(d_nbr, N) || !B.alloc(d_xy, 3 * (size_t)mesh->nnt) ||
        !B.alloc(d_cx, N) || !B.alloc(d_cy, N) || !B.alloc(d_ext, 4) || !B.alloc(d_key, N) || !B.alloc(d_skey, N) ||
        !B.alloc(d_ioflag, N) || !B.alloc(d_always, (size_t)ntiles) || !B.alloc(d_err, 1) || !B.alloc(d_tk, 3 * N) ||
        !B.alloc(d_stk, 3 * N) || !B.alloc(d_uk, 3 * N) || !B.alloc(d_adj, 3 * N) ||
        !B.alloc(d_off, (size_t)ntiles + 1) || !B.alloc(d_nu, 1))
        return CF_ERR_NOMEM;
    const u64 no_err = ~0ull;
    if (!up(d_conn, mesh->conn, 3 * N, s) || !up(d_bnd, mesh->boundary, 6 * N, s) || !up(d_err, &no_err, 1, s) ||
        hipMemsetAsync(d_always, 0, (size_t)ntiles, s) != hipSuccess)
        return CF_ERR_HIP;
    LongEntree le;
    for (int k = 0; k < CF_MAX_IO; k++) le.v[k] = params->long_entree[k];
    hipLaunchKernelGGL(nb0_kernel, dim3(blocks(N)), dim3(256), 0, s, d_conn, d_bnd, n, params->inlet, le, d_nb0, d_err);
    u64 err = 0;
    if (!down(&err, d_err, 1, s) || hipStreamSynchronize(s) != hipSuccess) return CF_ERR_HIP;
    if (err != ~0ull) return (err & 15) == 2 ? CF_ERR_ARG : CF_ERR_MESH;

    // internal order
    if (!keep_order && n > 1) {
        if (!up(d_xy, mesh->coords, 3 * (size_t)mesh->nnt, s)) return CF_ERR_HIP;
        hipLaunchKernelGGL(centroid_kernel, dim3(blocks(N)), dim3(256), 0, s, d_conn, d_xy, n, d_cx, d_cy);
        size_t bytes = 0, b2 = 0;
        GCHK(hipcub::DeviceReduce::Min(nullptr, bytes, d_cx, d_ext, n, s));
        GCHK(hipcub::DeviceReduce::Max(nullptr, b2, d_cx, d_ext, n, s));
        bytes = std::max(bytes, b2);
        char *tmp;
        if (!B.alloc(tmp, bytes)) return CF_ERR_NOMEM;
        GCHK(hipcub::DeviceReduce::Min(tmp, bytes, d_cx, d_ext + 0, n, s));
        GCHK(hipcub::DeviceReduce::Max(tmp, bytes, d_cx, d_ext + 1, n, s));
        GCHK(hipcub::DeviceReduce::Min(tmp, bytes, d_cy, d_ext + 2, n, s));
        GCHK(hipcub::DeviceReduce::Max(tmp, bytes, d_cy, d_ext + 3, n, s));
        hipLaunchKernelGGL(hilbert_key_kernel, dim3(blocks(N)), dim3(256), 0, s, d_cx, d_cy, n, d_ext, d_key, d_idx);
        int rc = sort_pairs(d_key, d_skey, d_idx, d_perm, n, 32, s, B);
        if (rc) return rc;
    } else {
        hipLaunchKernelGGL(iota_kernel, dim3(blocks(N)), dim3(256), 0, s, d_perm, n);
    }
    hipLaunchKernelGGL(inverse_kernel, dim3(blocks(N)), dim3(256), 0, s, d_perm, n, d_inv);
    hipLaunchKernelGGL(permute_kernel, dim3(blocks(N)), dim3(256), 0, s, d_nb0, d_conn, d_perm, d_inv, n, tile, d_nbr,
                       d_conn2, d_ioflag, d_always);

    // inflow/outflow cells: caller order, internal indices
    {
        size_t bytes = 0;
        hipcub::CountingInputIterator<int> cnt(0);
        GCHK(hipcub::DeviceSelect::Flagged(nullptr, bytes, cnt, d_ioflag, d_sel, d_nsel, n, s));
        char *tmp;
        if (!B.alloc(tmp, bytes)) return CF_ERR_NOMEM;
        GCHK(hipcub::DeviceSelect::Flagged(tmp, bytes, cnt, d_ioflag, d_sel, d_nsel, n, s));
    }
    // tile adjacency: sorted unique (tile, neighbour tile) keys
    int nu = 0;
    {
        const int M = (int)(3 * N);
        hipLaunchKernelGGL(tile_pair_kernel, dim3(blocks((size_t)M)), dim3(256), 0, s, d_nbr, n, tile, d_tk);
        size_t bytes = 0, b2 = 0;
        GCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, d_tk, d_stk, M, 0, 64, s));
        GCHK(hipcub::DeviceSelect::Unique(nullptr, b2, d_stk, d_uk, d_nu, M, s));
        bytes = std::max(bytes, b2);
        char *tmp;
        if (!B.alloc(tmp, bytes)) return CF_ERR_NOMEM;
        GCHK(hipcub::DeviceRadixSort::SortKeys(tmp, bytes, d_tk, d_stk, M, 0, 64, s));
        GCHK(hipcub::DeviceSelect::Unique(tmp, bytes, d_stk, d_uk, d_nu, M, s));
        if (!down(&nu, d_nu, 1, s) || hipStreamSynchronize(s) != hipSuccess) return CF_ERR_HIP;
        u64 last = 0;
        if (nu > 0 && (!down(&last, d_uk + (nu - 1), 1, s) || hipStreamSynchronize(s) != hipSuccess))
            return CF_ERR_HIP;
        if (nu > 0 && last == ~0ull) nu--;  // the key of the faces that do not cross a tile edge
        hipLaunchKernelGGL(tile_off_kernel, dim3(blocks((size_t)std::max(nu, ntiles + 1))), dim3(256), 0, s, d_uk, nu,
                           ntiles, d_off, d_adj);
    }
    int nsel = 0;
    if (!down(&nsel, d_nsel, 1, s) || hipStreamSynchronize(s) != hipSuccess) return CF_ERR_HIP;
    if (nsel > 0) hipLaunchKernelGGL(map_inv_kernel, dim3(blocks((size_t)nsel)), dim3(256), 0, s, d_sel, nsel, d_inv, d_io);
    GCHK(hipGetLastError());

    out.perm.resize(N);
    out.inv.resize(N);
    out.nbr.resize(N);
    out.conn.resize(3 * N);
    out.io_cells.resize((size_t)nsel);
    out.tile_off.resize((size_t)ntiles + 1);
    out.tile_adj.resize((size_t)nu);
    out.tile_always.resize((size_t)ntiles);
    static_assert(sizeof(Nbr) == sizeof(int4), "Nbr is four ints");
    if (!down(out.perm.data(), d_perm, N, s) || !down(out.inv.data(), d_inv, N, s) ||
        !down(reinterpret_cast<int4 *>(out.nbr.data()), d_nbr, N, s) || !down(out.conn.data(), d_conn2, 3 * N, s) ||
        !down(out.io_cells.data(), d_io, (size_t)nsel, s) || !down(out.tile_off.data(), d_off, (size_t)ntiles + 1, s) ||
        !down(out.tile_adj.data(), d_adj, (size_t)nu, s) || !down(out.tile_always.data(), d_always, (size_t)ntiles, s) ||
        hipStreamSynchronize(s) != hipSuccess)
        return CF_ERR_HIP;
    return CF_OK;
}

}  // namespace cfprep
