// wave_reduce.hpp -- wavefront-level reductions for gfx950 (wave64) with DPP.
//
// A wave's 64 lanes are reduced in six data-parallel-primitive moves and one
// readlane, with no LDS traffic and no barrier:
//   row_shr:1, 2, 4, 8      inclusive sums inside each row of 16 lanes
//                           (lane 15 of a row then holds the row's total)
//   row_bcast:15 (rows 1,3) lane 15 of a row added into the next row
//   row_bcast:31 (rows 2,3) lane 31 added into the upper half
// after which lane 63 holds the wave's total, read into a scalar register.
// Lanes whose DPP source lies outside the row, and rows the row mask leaves
// out, receive the operation's identity, so every step is a plain combine.
// The order of the combines is fixed, so a reduction is deterministic.
//
// Doubles move as two 32-bit halves.  Sums are carried as double-double
// (hi, lo) pairs combined with an error-free two-sum, so a reduction of n
// terms is within about n * 2^-106 of the exact sum of its terms: the
// rounded result equals the correctly rounded sum except for sums that
// cancel almost completely.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hydro {

constexpr int kDppRowShr1 = 0x111;
constexpr int kDppRowShr2 = 0x112;
constexpr int kDppRowShr4 = 0x114;
constexpr int kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142;
constexpr int kDppRowBcast31 = 0x143;

// x moved by DPP control CTRL; lanes without a source (or outside ROW_MASK)
// get `id`.
template <int CTRL, int ROW_MASK> __device__ __forceinline__ uint32_t dpp_u32(uint32_t x, uint32_t id) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)x, CTRL, ROW_MASK, 0xf, false);
}

template <int CTRL, int ROW_MASK> __device__ __forceinline__ double dpp_f64(double x, double id) {
    const uint64_t v = __builtin_bit_cast(uint64_t, x), i = __builtin_bit_cast(uint64_t, id);
    const uint32_t lo = dpp_u32<CTRL, ROW_MASK>((uint32_t)v, (uint32_t)i);
    const uint32_t hi = dpp_u32<CTRL, ROW_MASK>((uint32_t)(v >> 32), (uint32_t)(i >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ double readlane_f64(double x, int lane) {
    const uint64_t v = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Double-double value hi + lo with |lo| <= ulp(hi) / 2.
struct DD {
    double hi, lo;
};

// Error-free two-sum (Knuth): s + e == a + b exactly.
__host__ __device__ __forceinline__ DD two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return DD{s, e};
}

__host__ __device__ __forceinline__ DD dd_add(DD a, DD b) {
    DD s = two_sum(a.hi, b.hi);
    s.lo += a.lo + b.lo;
    return two_sum(s.hi, s.lo);
}

__host__ __device__ __forceinline__ DD dd_add(DD a, double b) {
    DD s = two_sum(a.hi, b);
    s.lo += a.lo;
    return two_sum(s.hi, s.lo);
}

// One reduction step of a DD sum: combine with the value DPP brings in.
template <int CTRL, int ROW_MASK> __device__ __forceinline__ DD dd_step(DD v) {
    const DD t{dpp_f64<CTRL, ROW_MASK>(v.hi, 0.0), dpp_f64<CTRL, ROW_MASK>(v.lo, 0.0)};
    return dd_add(v, t);
}

// Sum over the wave; every lane receives it.
__device__ __forceinline__ DD wave_sum(DD v) {
    v = dd_step<kDppRowShr1, 0xf>(v);
    v = dd_step<kDppRowShr2, 0xf>(v);
    v = dd_step<kDppRowShr4, 0xf>(v);
    v = dd_step<kDppRowShr8, 0xf>(v);
    v = dd_step<kDppRowBcast15, 0xa>(v);
    v = dd_step<kDppRowBcast31, 0xc>(v);
    return DD{readlane_f64(v.hi, 63), readlane_f64(v.lo, 63)};
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    v += dpp_u32<kDppRowShr1, 0xf>(v, 0);
    v += dpp_u32<kDppRowShr2, 0xf>(v, 0);
    v += dpp_u32<kDppRowShr4, 0xf>(v, 0);
    v += dpp_u32<kDppRowShr8, 0xf>(v, 0);
    v += dpp_u32<kDppRowBcast15, 0xa>(v, 0);
    v += dpp_u32<kDppRowBcast31, 0xc>(v, 0);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Minimum / maximum over the wave (fmin / fmax: a NaN lane is ignored).
__device__ __forceinline__ double wave_min(double v) {
    const double id = __builtin_huge_val();
    v = fmin(v, dpp_f64<kDppRowShr1, 0xf>(v, id));
    v = fmin(v, dpp_f64<kDppRowShr2, 0xf>(v, id));
    v = fmin(v, dpp_f64<kDppRowShr4, 0xf>(v, id));
    v = fmin(v, dpp_f64<kDppRowShr8, 0xf>(v, id));
    v = fmin(v, dpp_f64<kDppRowBcast15, 0xa>(v, id));
    v = fmin(v, dpp_f64<kDppRowBcast31, 0xc>(v, id));
    return readlane_f64(v, 63);
}

__device__ __forceinline__ double wave_max(double v) {
    const double id = -__builtin_huge_val();
    v = fmax(v, dpp_f64<kDppRowShr1, 0xf>(v, id));
    v = fmax(v, dpp_f64<kDppRowShr2, 0xf>(v, id));
    v = fmax(v, dpp_f64<kDppRowShr4, 0xf>(v, id));
    v = fmax(v, dpp_f64<kDppRowShr8, 0xf>(v, id));
    v = fmax(v, dpp_f64<kDppRowBcast15, 0xa>(v, id));
    v = fmax(v, dpp_f64<kDppRowBcast31, 0xc>(v, id));
    return readlane_f64(v, 63);
}

}  // namespace hydro
