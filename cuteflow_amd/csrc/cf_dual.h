// Forward-mode dual number (value + one tangent) for the tangent-linear
// model of the time step.  The value part of every operation is the same
// IEEE operation as on plain doubles, so a step evaluated on duals produces
// the forward step's bits in .v and the directional derivative in .d.
// Branches compare values only: the wet/dry switches, limiter regions and
// wave-speed cases are taken as in the forward (frozen).
#pragma once
#include "cf_hd.h"

namespace cf {

struct Dual {
    double v, d;
    CF_HD Dual() = default;
    CF_HD constexpr Dual(double x) : v(x), d(0.0) {}
    CF_HD constexpr Dual(double x, double dx) : v(x), d(dx) {}
};

CF_HD inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
CF_HD inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
CF_HD inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
CF_HD inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
CF_HD inline Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
CF_HD inline Dual &operator+=(Dual &a, Dual b) { return a = a + b; }
CF_HD inline Dual &operator-=(Dual &a, Dual b) { return a = a - b; }

CF_HD inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
CF_HD inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
CF_HD inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
CF_HD inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
CF_HD inline bool operator==(Dual a, Dual b) { return a.v == b.v; }
CF_HD inline bool operator!=(Dual a, Dual b) { return a.v != b.v; }
CF_HD inline bool operator<(Dual a, double b) { return a.v < b; }
CF_HD inline bool operator>(Dual a, double b) { return a.v > b; }
CF_HD inline bool operator<=(Dual a, double b) { return a.v <= b; }
CF_HD inline bool operator>=(Dual a, double b) { return a.v >= b; }
CF_HD inline bool operator<(double a, Dual b) { return a < b.v; }
CF_HD inline bool operator>(double a, Dual b) { return a > b.v; }

CF_HD inline double val(double x) { return x; }
CF_HD inline double val(Dual x) { return x.v; }

CF_HD inline double vsqrt(double x) { return sqrt(x); }
// d sqrt(x) = dx / (2 sqrt(x)); at x = 0 the tangent is 0: a zero depth
// there is a clamped (frozen) value whose own tangent is 0.
CF_HD inline Dual vsqrt(Dual x) {
    const double s = sqrt(x.v);
    return Dual(s, s > 0.0 ? x.d / (2.0 * s) : 0.0);
}
CF_HD inline double vabs(double x) { return fabs(x); }
CF_HD inline Dual vabs(Dual x) { return Dual(fabs(x.v), x.v < 0.0 ? -x.d : x.d); }

}  // namespace cf
