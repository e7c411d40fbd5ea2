// Host-side mesh preprocessing: directed-edge neighbour matching and boundary
// classification (the reference's tab_bound_tri + boundary_kind), inlet
// lengths, the internal cell order and tile adjacency of the HIP step.
#include "cf_prep.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

struct Edge {
    int64_t key;
    int cell;
};

inline int64_t ekey(int a, int b) { return ((int64_t)a << 32) | (uint32_t)b; }

// Index along a Hilbert curve of order 16 on a 65536 x 65536 grid (the
// classic xy -> d conversion with quadrant rotation).
uint32_t hilbert_d(uint32_t x, uint32_t y) {
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
            std::swap(x, y);
        }
    }
    return d;
}

}  // namespace

// Neighbours by the reversed directed edge (maillage.cuf:881-914, all-pairs in
// the reference; here a sort of the 3N directed edges), then the boundary
// classes of the faces left without a neighbour (maillage.cuf:981-1110):
// inlet (-1) when both nodes are inlet nodes, else outlet (-2) when both are
// outlet nodes, else wall (-3) when both are wall nodes, else wall when one
// node is a wall node and the other an inlet/outlet node.
extern "C" int cf_build_boundary(int nelt, int nnt, const int *conn, int ndi, const int *ndinput,
                                 const int *num_in, int ndo, const int *ndoutput, const int *num_out,
                                 int ndw, const int *ndwall, int *boundary) {
    if (nelt <= 0 || nnt <= 0 || !conn || !boundary || ndi < 0 || ndo < 0 || ndw < 0 ||
        (ndi > 0 && !ndinput) || (ndo > 0 && !ndoutput) || (ndw > 0 && !ndwall))
        return CF_ERR_ARG;
    const size_t N = (size_t)nelt;
    for (size_t i = 0; i < 3 * N; i++)
        if (conn[i] < 1 || conn[i] > nnt) return CF_ERR_MESH;
    std::vector<Edge> ed(3 * N);
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) {
            ed[3 * i + k].key = ekey(conn[3 * i + k], conn[3 * i + (k + 1) % 3]);
            ed[3 * i + k].cell = (int)i;
        }
    std::sort(ed.begin(), ed.end(), [](const Edge &x, const Edge &y) {
        return x.key != y.key ? x.key < y.key : x.cell < y.cell;
    });
    // boundary(i,k,1) at [k*N + i], boundary(i,k,2) at [3N + k*N + i]
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) {
            const int a = conn[3 * i + k], b = conn[3 * i + (k + 1) % 3];
            const int64_t want = ekey(b, a);
            auto it = std::lower_bound(ed.begin(), ed.end(), want,
                                       [](const Edge &x, int64_t v) { return x.key < v; });
            int nb = (int)i + 1;
            for (; it != ed.end() && it->key == want; ++it)
                if (it->cell != (int)i) nb = it->cell + 1;  // highest matching id
            boundary[k * N + i] = nb;
            boundary[3 * N + k * N + i] = (int)i + 1;
        }
    // first position of each node in each list
    auto positions = [&](int nd, const int *list) {
        std::vector<int> pos((size_t)nnt + 1, -1);
        for (int q = nd - 1; q >= 0; q--)
            if (list[q] >= 1 && list[q] <= nnt) pos[list[q]] = q;
        return pos;
    };
    std::vector<int> pi = positions(ndi, ndinput), po = positions(ndo, ndoutput), pw = positions(ndw, ndwall);
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) {
            if (boundary[k * N + i] != (int)i + 1) continue;
            const int n1 = conn[3 * i + k], n2 = conn[3 * i + (k + 1) % 3];
            if (ndi > 0 && n1 != n2 && pi[n1] >= 0 && pi[n2] >= 0) {
                const int q = std::min(pi[n1], pi[n2]);
                boundary[k * N + i] = -1;
                boundary[3 * N + k * N + i] = num_in ? num_in[q] : 1;
                continue;
            }
            if (ndo > 0 && n1 != n2 && po[n1] >= 0 && po[n2] >= 0) {
                const int q = std::min(po[n1], po[n2]);
                boundary[k * N + i] = -2;
                boundary[3 * N + k * N + i] = num_out ? num_out[q] : 1;
                continue;
            }
            if (ndw > 0 && n1 != n2 && pw[n1] >= 0 && pw[n2] >= 0) {
                boundary[k * N + i] = -3;
                continue;
            }
            if (ndi > 0 && ndw > 0 && ((pw[n1] >= 0 && pi[n2] >= 0) || (pi[n1] >= 0 && pw[n2] >= 0)))
                boundary[k * N + i] = -3;
            if (ndo > 0 && ndw > 0 && ((pw[n1] >= 0 && po[n2] >= 0) || (po[n1] >= 0 && pw[n2] >= 0)))
                boundary[k * N + i] = -3;
        }
    return CF_OK;
}

// calcul_long_entree (maillage.cuf:1423-1452)
extern "C" int cf_inlet_lengths(int nelt, int nnt, const int *conn, const double *coords,
                                const int *boundary, int nentrees, double *long_entree) {
    if (nelt <= 0 || !conn || !coords || !boundary || !long_entree || nentrees < 0 || nentrees > CF_MAX_IO)
        return CF_ERR_ARG;
    const size_t N = (size_t)nelt;
    for (int k = 0; k < nentrees; k++) long_entree[k] = 0.0;
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) {
            if (boundary[k * N + i] != -1) continue;
            const int id = boundary[3 * N + k * N + i];
            const int n1 = conn[3 * i + k] - 1, n2 = conn[3 * i + (k + 1) % 3] - 1;
            if (id < 1 || id > nentrees || n1 < 0 || n2 < 0 || n1 >= nnt || n2 >= nnt) continue;
            const double dx = coords[3 * n1] - coords[3 * n2], dy = coords[3 * n1 + 1] - coords[3 * n2 + 1];
            long_entree[id - 1] = long_entree[id - 1] + std::sqrt(dx * dx + dy * dy);
        }
    return CF_OK;
}

namespace cfprep {

int prepare(const cf_mesh *mesh, const cf_params *params, int tile, bool keep_order, Prep &out) {
    const int nelt = mesh->nelt;
    const size_t N = (size_t)nelt;
    const int *conn = mesh->conn;
    const int *b = mesh->boundary;
    out.nelt = nelt;
    out.tile = tile;

    // caller-order neighbour codes (set_undefined_boudaries_as_walls applied)
    std::vector<Nbr> nb0(N);
    for (size_t i = 0; i < N; i++) {
        for (int k = 0; k < 3; k++) {
            int v = b[k * N + i];
            const int id = b[3 * N + k * N + i];
            if (v == (int)i + 1) v = -3;
            if (v > 0) {
                if (v > nelt) return CF_ERR_MESH;
                nb0[i].c[k] = v - 1;
            } else if (v == -1 || v == -2) {
                if (id < 1 || id > CF_MAX_IO) return CF_ERR_MESH;
                if (v == -1 && params->inlet == 2 && !(params->long_entree[id - 1] > 0.0)) return CF_ERR_ARG;
                nb0[i].c[k] = -((v == -1 ? 1 : 2) | (id << 2));
            } else if (v == -3) {
                nb0[i].c[k] = -3;
            } else {
                return CF_ERR_MESH;
            }
        }
        // face of the neighbour sharing the edge (flux_riemann.cuf:262-277);
        // boundary faces point at the cell itself, same face
        int w = 0;
        for (int j = 0; j < 3; j++) {
            int tjv = j;
            const int c = nb0[i].c[j];
            if (c >= 0) {
                const int A = conn[3 * i + j], B = conn[3 * i + (j + 1) % 3];
                const int *q = conn + 3 * (size_t)c;
                if ((q[0] == A && q[1] == B) || (q[1] == A && q[0] == B)) tjv = 0;
                else if ((q[1] == A && q[2] == B) || (q[2] == A && q[1] == B)) tjv = 1;
                else if ((q[2] == A && q[0] == B) || (q[0] == A && q[2] == B)) tjv = 2;
                else tjv = 0;  // no face matches (non-conforming mesh): the first face
            }
            w |= tjv << (2 * j);
        }
        nb0[i].w = w;
    }

    // internal order
    out.perm.resize(N);
    for (size_t i = 0; i < N; i++) out.perm[i] = (int)i;
    if (!keep_order && N > 1) {
        const double *xy = mesh->coords;
        std::vector<double> cx(N), cy(N);
        double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
        for (size_t i = 0; i < N; i++) {
            double sx = 0.0, sy = 0.0;
            for (int k = 0; k < 3; k++) {
                sx += xy[3 * (size_t)(conn[3 * i + k] - 1)];
                sy += xy[3 * (size_t)(conn[3 * i + k] - 1) + 1];
            }
            cx[i] = sx / 3.0;
            cy[i] = sy / 3.0;
            x0 = std::min(x0, cx[i]); x1 = std::max(x1, cx[i]);
            y0 = std::min(y0, cy[i]); y1 = std::max(y1, cy[i]);
        }
        // one scale for both axes keeps patches compact on elongated domains
        const double span = std::max(std::max(x1 - x0, y1 - y0), 1e-300);
        const double sc = 65535.0 / span;
        std::vector<std::pair<uint32_t, int>> key(N);
        for (size_t i = 0; i < N; i++) {
            const uint32_t qx = (uint32_t)std::min(65535.0, std::max(0.0, (cx[i] - x0) * sc));
            const uint32_t qy = (uint32_t)std::min(65535.0, std::max(0.0, (cy[i] - y0) * sc));
            key[i] = {hilbert_d(qx, qy), (int)i};
        }
        std::sort(key.begin(), key.end());
        for (size_t i = 0; i < N; i++) out.perm[i] = key[i].second;
    }
    out.inv.resize(N);
    for (size_t i = 0; i < N; i++) out.inv[out.perm[i]] = (int)i;

    out.nbr.resize(N);
    out.conn.resize(3 * N);
    for (size_t n = 0; n < N; n++) {
        const int o = out.perm[n];
        Nbr r = nb0[o];
        for (int k = 0; k < 3; k++)
            if (r.c[k] >= 0) r.c[k] = out.inv[r.c[k]];
        out.nbr[n] = r;
        for (int k = 0; k < 3; k++) out.conn[3 * n + k] = conn[3 * (size_t)o + k];
    }

    // cells with an inflow or outflow face, in the caller's order (the order
    // in which their discharges are summed)
    out.io_cells.clear();
    for (size_t o = 0; o < N; o++)
        for (int k = 0; k < 3; k++) {
            const int c = nb0[o].c[k];
            if (c < 0 && (((-c) & 3) == 1 || ((-c) & 3) == 2)) {
                out.io_cells.push_back(out.inv[o]);
                break;
            }
        }

    // tiles of `tile` consecutive internal cells: neighbour tiles and the
    // tiles that must always run (inflow faces give a flux into dry cells)
    const int ntiles = (int)((N + tile - 1) / tile);
    out.tile_off.assign(ntiles + 1, 0);
    out.tile_adj.clear();
    out.tile_always.assign(ntiles, 0);
    std::vector<int> tmp;
    for (int t = 0; t < ntiles; t++) {
        tmp.clear();
        const size_t a = (size_t)t * tile, e = std::min(N, a + tile);
        for (size_t n = a; n < e; n++)
            for (int k = 0; k < 3; k++) {
                const int c = out.nbr[n].c[k];
                if (c >= 0) {
                    const int tt = c / tile;
                    if (tt != t) tmp.push_back(tt);
                } else if (((-c) & 3) == 1) {
                    out.tile_always[t] = 1;
                }
            }
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        out.tile_adj.insert(out.tile_adj.end(), tmp.begin(), tmp.end());
        out.tile_off[t + 1] = (int)out.tile_adj.size();
    }
    return CF_OK;
}

}  // namespace cfprep
