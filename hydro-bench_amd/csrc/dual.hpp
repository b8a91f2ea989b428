// dual.hpp -- forward-mode dual number for the tangent-linear model.
//
// Dual carries a value and one directional derivative.  Every operation the
// timestep arithmetic uses is defined so that physics.hpp instantiated with
// Real = Dual computes, alongside the unchanged fp64 value (same operations in
// the same order), the derivative of that value along one input direction.
// Comparisons look at the value only, so every branch is taken exactly as in
// the fp64 forward step; min/max/abs/copysign differentiate the branch taken.
#pragma once
#include <hip/hip_runtime.h>

#include "pow_fixed.hpp"

namespace hydro {
namespace ad {

// Lives in its own namespace so that unqualified math calls on double/float
// in namespace hydro never see these overloads; Dual arguments find them by
// argument-dependent lookup.
struct Dual {
    double v, d;
    __host__ __device__ constexpr Dual() : v(0), d(0) {}
    __host__ __device__ constexpr Dual(double x) : v(x), d(0) {}
    __host__ __device__ constexpr Dual(double x, double dx) : v(x), d(dx) {}
    __host__ __device__ explicit operator double() const { return v; }
};

__host__ __device__ inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__host__ __device__ inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__host__ __device__ inline Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ inline Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
__host__ __device__ inline Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
__host__ __device__ inline Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
__host__ __device__ inline Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
__host__ __device__ inline Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
__host__ __device__ inline Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
__host__ __device__ inline Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
__host__ __device__ inline Dual operator/(double a, Dual b) {
    const double q = a / b.v;
    return Dual(q, -q * b.d / b.v);
}
__host__ __device__ inline Dual &operator+=(Dual &a, Dual b) { return a = a + b; }
__host__ __device__ inline Dual &operator-=(Dual &a, Dual b) { return a = a - b; }

#define HYDRO_DUAL_CMP(op)                                                                   \
    __host__ __device__ inline bool operator op(Dual a, Dual b) { return a.v op b.v; }       \
    __host__ __device__ inline bool operator op(Dual a, double b) { return a.v op b; }       \
    __host__ __device__ inline bool operator op(double a, Dual b) { return a op b.v; }
HYDRO_DUAL_CMP(<)
HYDRO_DUAL_CMP(<=)
HYDRO_DUAL_CMP(>)
HYDRO_DUAL_CMP(>=)
HYDRO_DUAL_CMP(==)
HYDRO_DUAL_CMP(!=)
#undef HYDRO_DUAL_CMP

__host__ __device__ inline Dual sqrt(Dual a) {
    const double r = ::sqrt(a.v);
    return Dual(r, r > 0 ? a.d / (2.0 * r) : 0.0);
}
// Values come from the same library calls as the fp64 path (so signed zeros
// and NaN handling agree); derivatives follow the operand taken (the first
// on ties).
__host__ __device__ inline Dual fabs(Dual a) { return Dual(::fabs(a.v), a.v < 0 ? -a.d : a.d); }
__host__ __device__ inline Dual fmax(Dual a, Dual b) {
    return Dual(::fmax(a.v, b.v), (a.v >= b.v || b.v != b.v) ? a.d : b.d);
}
__host__ __device__ inline Dual fmin(Dual a, Dual b) {
    return Dual(::fmin(a.v, b.v), (a.v <= b.v || b.v != b.v) ? a.d : b.d);
}
__host__ __device__ inline Dual fmax(Dual a, double b) { return fmax(a, Dual(b)); }
__host__ __device__ inline Dual fmin(Dual a, double b) { return fmin(a, Dual(b)); }
__host__ __device__ inline Dual fmin(double a, Dual b) { return fmin(Dual(a), b); }
__host__ __device__ inline Dual copysign(Dual a, Dual s) {
    const bool flip = __builtin_signbit(a.v) != __builtin_signbit(s.v);
    return Dual(::copysign(a.v, s.v), flip ? -a.d : a.d);
}
__host__ __device__ inline Dual copysign(Dual a, double s) { return copysign(a, Dual(s)); }
__host__ __device__ inline Dual copysign(double a, Dual s) { return copysign(Dual(a), s); }
// x^y for constant y
__host__ __device__ inline Dual pow_d(Dual x, double y) {
    const double p = hydro::pow_d(x.v, y);
    return Dual(p, x.v != 0 ? y * p / x.v * x.d : 0.0);
}
__host__ __device__ inline Dual pow_rare(Dual x, double y) { return pow_d(x, y); }

}  // namespace ad
}  // namespace hydro
