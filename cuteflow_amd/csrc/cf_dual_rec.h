// Dual number whose comparisons also record their outcome, for diagnostic
// tangent-linear runs on the GPU: the arithmetic is Dual's, operation for
// operation (so values and tangents are Dual's bits), and every comparison
// of two values -- the wet/dry switches, dry-face thresholds, wave-speed
// cases, solver regions, max/min selections -- mixes its result into the
// calling thread's branch signature in LDS.  Two runs whose signatures of a
// (step, cell) differ took some branch of that cell's update differently.
#pragma once
#include <hip/hip_runtime.h>

#include "cf_device.h"
#include "cf_dual.h"

namespace cf {

constexpr int kSigThreads = 256;  // largest workgroup of the per-cell kernels
__shared__ unsigned s_branch_sig[kSigThreads];

__device__ inline bool branch_rec(bool r) {
    unsigned &s = s_branch_sig[threadIdx.x];
    s = (s * 0x01000193u) ^ (r ? 0x9e3779b9u : 0x7f4a7c15u);
    return r;
}

struct DualR {
    double v, d;
    __device__ DualR() = default;
    __device__ constexpr DualR(double x) : v(x), d(0.0) {}
    __device__ constexpr DualR(double x, double dx) : v(x), d(dx) {}
};

__device__ inline DualR operator-(DualR a) { return DualR(-a.v, -a.d); }
__device__ inline DualR operator+(DualR a, DualR b) { return DualR(a.v + b.v, a.d + b.d); }
__device__ inline DualR operator-(DualR a, DualR b) { return DualR(a.v - b.v, a.d - b.d); }
__device__ inline DualR operator*(DualR a, DualR b) { return DualR(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ inline DualR operator/(DualR a, DualR b) {
    const double q = a.v / b.v;
    return DualR(q, (a.d - q * b.d) / b.v);
}
__device__ inline DualR &operator+=(DualR &a, DualR b) { return a = a + b; }
__device__ inline DualR &operator-=(DualR &a, DualR b) { return a = a - b; }

__device__ inline bool operator<(DualR a, DualR b) { return branch_rec(a.v < b.v); }
__device__ inline bool operator>(DualR a, DualR b) { return branch_rec(a.v > b.v); }
__device__ inline bool operator<=(DualR a, DualR b) { return branch_rec(a.v <= b.v); }
__device__ inline bool operator>=(DualR a, DualR b) { return branch_rec(a.v >= b.v); }
__device__ inline bool operator==(DualR a, DualR b) { return branch_rec(a.v == b.v); }
__device__ inline bool operator!=(DualR a, DualR b) { return branch_rec(a.v != b.v); }
__device__ inline bool operator<(DualR a, double b) { return branch_rec(a.v < b); }
__device__ inline bool operator>(DualR a, double b) { return branch_rec(a.v > b); }
__device__ inline bool operator<=(DualR a, double b) { return branch_rec(a.v <= b); }
__device__ inline bool operator>=(DualR a, double b) { return branch_rec(a.v >= b); }
__device__ inline bool operator<(double a, DualR b) { return branch_rec(a < b.v); }
__device__ inline bool operator>(double a, DualR b) { return branch_rec(a > b.v); }

__device__ inline double val(DualR x) { return x.v; }
__device__ inline DualR vsqrt(DualR x) {
    const double s = sqrt(x.v);
    return DualR(s, s > 0.0 ? x.d / (2.0 * s) : 0.0);
}
__device__ inline DualR vabs(DualR x) { return DualR(fabs(x.v), x.v < 0.0 ? -x.d : x.d); }
__device__ inline DualR cf_pow(DualR x, double y) {
    const double p = cf_pow(x.v, y);
    return DualR(p, x.d * ((y * p) / x.v));
}

}  // namespace cf
