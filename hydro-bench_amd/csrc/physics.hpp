// physics.hpp -- per-slot flux and per-cell update of the hydro-cal timestep,
// device side, templated on Real (double / float) and on the variant
// (0 = A: cuda_native_impl/hydro-cal-src/src/functors.cu,
//  1 = B: cuda_native_impl/baselines/native_cuda_v0/fp64/hydro-cal-src/src/functors.cu).
//
// Arithmetic is written so that every rounding happens where the reference's
// does: Real-typed locals, double literals (so the float build promotes the
// same sub-expressions to double that `using Real = float` does in
// cuda_native_impl/fp32_src), sqrt_r = sqrt / sqrtf, pow in double.  The
// library is compiled with -ffp-contract=off so no FMA changes a rounding.
//
// Structure differs from the reference on purpose (MI355X-first):
//  * the flux of a slot is returned in registers; the caller decides whether
//    it is stored (two-kernel path) or consumed by the fused cell update;
//  * the primary-side and mirror-side Osher solves (F:151-200) are one
//    Riemann solve on (left, right) states chosen with selects, so a wave
//    with both kinds of slot runs the solver once, not twice;
//  * the 16-case Osher table (F:612-711) is evaluated as five predicated
//    terms in the table's fixed order (left state, left sonic point,
//    intermediate state, right sonic point, right state), each computed at
//    most once per lane and skipped by the whole wave when no lane needs it.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#include "pow_fixed.hpp"

namespace hydro {

// Slot kinds, one per slot of the reordered layout (low 4 bits of the slot
// meta byte).  The first two are interior slots that reach the Osher branch
// as the primary side (KLAS 0 and original cell id < original neighbour id,
// F:151) or as the mirror side (any other non-boundary KLAS, F:165).
enum SlotKind : int {
    K_PRIMARY = 0,
    K_MIRROR = 1,
    K_B10 = 2,   // discharge            (F:204-232)
    K_B1 = 3,    // water level          (F:371-392)
    K_B3 = 4,    // rating curve         (F:320-369)
    K_B4 = 5,    // wall                 (F:544-550)
    K_B5 = 6,    // open                 (F:551-557)
    K_B6 = 7,    // weir                 (F:258-318)
    K_B7 = 8,    // breach               (F:234-256)
    K_BUNK = 9,  // boundary code without a handler: only FLUX2 is written
    K_PAD = 15   // triangle padding slot (NAC 0, SIDE 0, KLAS 0, COSF SINF 0)
};
constexpr int META_KIND_MASK = 0x0f;
constexpr int META_PRIMARY_BIT = 0x10;  // kind to take if a KLAS 7 slot becomes 0

// KLAS -> kind (F:109 decides boundary vs interior; F:151 primary vs mirror).
__host__ __device__ inline int kind_of_klas(int klas, bool primary_side) {
    if ((klas >= 1 && klas <= 8) || klas >= 10) {
        switch (klas) {
        case 10: return K_B10;
        case 1: return K_B1;
        case 3: return K_B3;
        case 4: return K_B4;
        case 5: return K_B5;
        case 6: return K_B6;
        case 7: return K_B7;
        default: return K_BUNK;
        }
    }
    return (klas == 0 && primary_side) ? K_PRIMARY : K_MIRROR;
}

// The float overloads of the math functions, on the host as on the device
// (an unqualified fmin(float, float) would otherwise find ::fmin(double,
// double) in a host build and evaluate the fp32 tree's expressions in double).
using std::copysign;
using std::fabs;
using std::fmax;
using std::fmin;
using std::sqrt;

template <class R> struct Consts {
    static constexpr R C0 = (R)1.33;
    static constexpr R C1 = (R)1.7;
    static constexpr R VMIN = (R)0.001;
    static constexpr R QLUA = (R)0.0;
    static constexpr R BRDTH = (R)100.0;
};

// sqrt of the fp64 tree / sqrtf of the fp32 tree (argument rounded to float
// first, as sqrtf(double) does); for the dual number, its own sqrt.
template <class R> struct SqrtR {
    template <class X> __host__ __device__ static __forceinline__ R f(X x) { return sqrt(x); }
};
template <> struct SqrtR<float> {
    __host__ __device__ static __forceinline__ float f(double x) { return sqrtf((float)x); }
};
template <> struct SqrtR<double> {
    __host__ __device__ static __forceinline__ double f(double x) { return sqrt(x); }
};
template <class R, class X> __host__ __device__ __forceinline__ R sqrt_r(X x) { return SqrtR<R>::f(x); }
// pow_d / pow_rare (pow in double with the timestep's fixed exponents) come
// from pow_fixed.hpp; the dual number's overloads are found by argument-
// dependent lookup.

template <class R> struct Flux { R f0, f1, f2, f3; };

// Mesh-wide scalars as the kernels use them.
template <class R> struct Scalars {
    R DT, HM1, HM2, JL;
};

// Element type of the arrays a Real of the arithmetic type R is read from
// (the dual number of the tangent-linear model reads double arrays).
template <class R> struct StorageOf {
    using type = R;
};

// Everything the boundary handlers may read besides the cell states.
template <class R> struct BcCtx {
    using T = typename StorageOf<R>::type;
    const T *QT, *DQT, *ZT, *DZT, *ZW, *QW, *BF;
    const int32_t *NHQ1;
    int CELL, NHQ;
    int cell;   // ORIGINAL cell id (rows of the dense boundary arrays)
    int kt, jt;
    R SIDE, COSF, SINF;
};

// ---------------------------------------------------------------------------
// Osher approximate Riemann flux (F:570-712 with QS F:445-503, QF F:505-511).
template <class R>
__host__ __device__ __forceinline__ void qf_acc(R acc[4], R h, R u, R v, bool neg) {
    R f0 = h * u;
    R f1 = f0 * u;
    R f2 = f0 * v;
    R f3 = 4.905 * h * h;
    if (neg) { f0 = -f0; f1 = -f1; f2 = -f2; f3 = -f3; }
    acc[0] = acc[0] + f0;
    acc[1] = acc[1] + f1;
    acc[2] = acc[2] + f2;
    acc[3] = acc[3] + f3;
}

template <class R>
__host__ __device__ __forceinline__ void osher(R H_pre, R QL0, R QL1, R QL2, R QR0, R QR1, R QR2,
                                               R FIL, R out[4]) {
    const R CR = sqrt_r<R>(9.81 * QR0);
    const R FIR = QR1 - 2 * CR;
    const R UA = (FIL + FIR) / 2;
    const R CA = fabs((FIL - FIR) / 4);
    const R CL = sqrt_r<R>(9.81 * H_pre);
    // K2 from the intermediate state, K1 from the end states (F:586-610;
    // the cases the reference leaves undefined take 4, as in the Taichi port).
    int K2 = (CA < UA) ? 1 : (UA >= 0.0 && UA < CA) ? 2 : (UA >= -CA && UA < 0.0) ? 3 : 4;
    const bool rsub = (QR1 >= -CR);
    int K1 = (QL1 < CL && rsub) ? 1 : (QL1 >= CL && rsub) ? 2 : (QL1 < CL && QR1 < -CR) ? 3 : 4;
    // K1 in {2, 4}: the flux starts from the left state; K1 in {3, 4}: it
    // ends at the right state.  Written as single bit tests: the gfx950
    // backend of this toolchain mis-lowers the two-case switch that
    // (K1 == 2 || K1 == 4) becomes, dropping the K1 == 2 negation of the
    // left sonic-point term (seen at K1 = 2, K2 = 2 with uniform operands).
    const bool lsup = (K1 & 1) == 0;
    const bool rneg = K1 >= 3;
    out[0] = 0; out[1] = 0; out[2] = 0; out[3] = 0;
    if (lsup) qf_acc<R>(out, QL0, QL1, QL2, false);
    if ((K2 == 1) != lsup) {                     // sonic point of the left curve
        R US = FIL / 3;
        R HS = US * US / 9.81;
        qf_acc<R>(out, HS, US, QL2, lsup);
    }
    R FIRm = FIR;
    if (K2 == 2 || K2 == 3) {                    // intermediate state
        const bool left = (K2 == 2);
        R X = (left ? FIL : FIR) - UA;
        if (!left) FIRm = X;                     // QS<5> updates FIR in place
        R HA = X * X / 39.24;
        qf_acc<R>(out, HA, UA, left ? QL2 : QR2, false);
    }
    if ((K2 == 4) != rneg) {                     // sonic point of the right curve
        R US = FIRm / 3;
        R HS = US * US / 9.81;
        qf_acc<R>(out, HS, US, QR2, rneg);
    }
    if (rneg) qf_acc<R>(out, QR0, QR1, QR2, false);
}

// ---------------------------------------------------------------------------
// QD (F:394-421) and LAQP (F:423-443).
template <class R> __host__ __device__ inline R qd(R ZL, R ZR, R ZB) {
    const R CM = 0.384, SIGMA = 0.667, FI = 4.43;
    R ZU = fmax(ZL, ZR), ZD = fmin(ZL, ZR);
    R H0 = ZU - ZB, HS = ZD - ZB;
    R DELTA = HS / H0;
    if (DELTA <= SIGMA) return copysign(CM * pow_rare(H0, 1.5), (double)(ZL - ZR));
    R DH = ZU - ZD;
    if (DH > 0.09) return copysign(FI * HS * sqrt_r<R>(DH), ZL - ZR);
    return copysign(FI * HS * 0.3 * DH / 0.1, (double)(ZL - ZR));
}

template <class R, class T> __host__ __device__ inline R laqp(R X, const T *A, const T *B, int MS) {
    if (MS <= 0) return 0;
    if (X < A[0]) return (R)B[0];
    if (X > A[MS - 1]) return (R)B[MS - 1];
    for (int i = 0; i < MS - 1; i++)
        if (X >= A[i] && X <= A[i + 1]) return B[i] + (B[i + 1] - B[i]) / (A[i + 1] - A[i]) * (X - A[i]);
    return 0;
}

template <class R> __host__ __device__ inline R urb_iterate(R QLn, R H_pre, R HB1, double tol) {
    R FIAL = QLn + 6.264 * sqrt_r<R>(H_pre);
    R UR0 = QLn;
    R URB = UR0;
    for (int it = 1; it <= 30; it++) {
        R FIAR = URB - 6.264 * sqrt_r<R>(HB1);
        URB = (FIAL + FIAR) * (FIAL - FIAR) * (FIAL - FIAR) / HB1 / 313.92;
        if (fabs(URB - UR0) <= tol) break;
        UR0 = URB;
    }
    return URB;
}

// ---------------------------------------------------------------------------
// Boundary slot (BOUNDA F:513-568).  `prev` supplies FLUX0/1/3 for codes
// that do not write them.  Sets *klas7_fired when a KLAS 7 slot turns into
// an interior slot (F:242).
// RARE = false compiles only the kinds the common meshes use (KLAS 10, 1, 4);
// the caller guarantees no other boundary kind occurs.
template <class R, int VAR, bool RARE = true>
__host__ __device__ inline Flux<R> boundary_flux(int kind, const Scalars<R> &S, const BcCtx<R> &bc,
                                                 R H1, R U1, R V1, R Z1, R ZBp, R ZC, R UC, R VC,
                                                 R QL1, R QL2, R FIL, Flux<R> prev,
                                                 bool *klas7_fired) {
    using K = Consts<R>;
    Flux<R> f;
    const R CL = sqrt_r<R>(9.81 * H1);
    const bool super = (VAR == 1) ? (QL1 > CL) : (QL1 > CL && H1 < S.HM2);
    if (super) {
        f.f0 = H1 * QL1;
        f.f1 = f.f0 * QL1;
        f.f2 = f.f0 * QL2;
        f.f3 = 4.905 * H1 * H1;
        return f;
    }
    f.f2 = 0;
    if (QL1 > 0) f.f2 = H1 * QL1 * QL2;
    f.f0 = prev.f0;
    f.f1 = prev.f1;
    f.f3 = prev.f3;
    const size_t row = (size_t)bc.jt * (size_t)bc.CELL + (size_t)bc.cell;
    switch (kind) {
    case K_B10: {
        R F0 = -(bc.QT[row] + bc.DQT[row] * bc.kt);
        F0 = F0 / bc.SIDE;
        R QB2 = F0 * F0;
        R HB0 = H1;
        R HB = 0;
        for (int k = 1; k <= 20; k++) {
            R Wt = FIL - F0 / HB0;
            HB = Wt * Wt / 39.24;
            if (fabs(HB0 - HB) <= 0.005) break;
            HB0 = HB0 * 0.5 + HB * 0.5;
        }
        f.f0 = F0;
        f.f1 = (HB <= 1) ? (R)0 : QB2 / HB;
        f.f2 = 0;
        f.f3 = 4.905 * HB * HB;
    } break;
    case K_B1: {
        R HB1 = bc.ZT[row] + bc.DZT[row] * (VAR == 1 ? bc.kt : bc.jt) - ZBp;
        HB1 = fmax(HB1, S.HM2);
        R URB = urb_iterate<R>(QL1, H1, HB1, 0.0001);
        f.f0 = HB1 * URB;
        f.f1 = f.f0 * URB;
        f.f3 = 4.905 * HB1 * HB1;
    } break;
    case K_B3: {
        if (!RARE) break;
        R QZH3 = (U1 * bc.COSF + V1 * bc.SINF) * H1 * bc.SIDE;
        QZH3 = fmax(QZH3, (R)0.0);
        const auto *QWr = bc.QW + (size_t)bc.cell * (size_t)bc.NHQ;
        const auto *ZWr = bc.ZW + (size_t)bc.cell * (size_t)bc.NHQ;
        R ZQH1 = laqp(QZH3, QWr, ZWr, bc.NHQ1[bc.cell]);
        R ZQH11 = ZQH1 - Z1;
        R HB1 = ZQH1 - ZBp;
        HB1 = fmax(HB1, S.HM2);
        R q = fmax(QL1, (R)0.0);
        if (QZH3 <= QWr[1] || HB1 <= S.HM2 || ZQH11 >= 0.1) {
            f.f0 = H1 * q;
            f.f1 = f.f0 * q;
            f.f3 = 4.905 * H1 * H1;
        } else {
            R URB = urb_iterate<R>(q, H1, HB1, 0.001);
            f.f0 = HB1 * URB;
            f.f1 = f.f0 * URB;
            f.f3 = 4.905 * HB1 * HB1;
        }
    } break;
    case K_B4:
        f.f0 = 0; f.f1 = 0; f.f2 = 0;
        f.f3 = 4.905 * H1 * H1;
        break;
    case K_B5: {
        if (!RARE) break;
        R q = fmax(QL1, (R)0.0);
        f.f0 = H1 * q;
        f.f1 = f.f0 * q;
        f.f3 = 4.905 * H1 * H1 * (1.0 - S.JL) * (1.0 - S.JL);
    } break;
    case K_B6: {
        if (!RARE) break;
        const R TOP = (R)bc.BF[bc.cell];
        const R cs = bc.COSF, sn = bc.SINF;
        if (Z1 <= TOP && ZC <= TOP) {
            f.f0 = 0; f.f1 = 0; f.f2 = 0;
            f.f3 = 4.905 * H1 * H1;
        } else if (Z1 > TOP && ZC <= TOP) {
            f.f0 = K::C0 * pow_rare(Z1 - TOP, 1.5);
            f.f1 = f.f0 * QL1;
            f.f2 = f.f0 * QL2;
            f.f3 = 4.905 * pow_rare(TOP - ZBp, 2.0);
        } else if (Z1 <= TOP && ZC > TOP) {
            f.f0 = -K::C0 * pow_rare(ZC - TOP, 1.5);
            f.f1 = f.f0 * fmin(UC * cs + VC * sn, (R)0.0);
            f.f2 = f.f0 * (VC * cs - UC * sn);
            f.f3 = 4.905 * pow_rare(Z1 - ZBp, 2.0);
        } else {
            R DZ = fabs(Z1 - ZC);
            if (Z1 <= ZC) {
                R HD = Z1 - TOP;
                R UN = fmin(UC * cs + VC * sn, (R)0.0);
                R VT = VC * cs - UC * sn;
                R SH = HD + DZ;
                R CE = fmin(1.0, 1.05 * pow_rare(DZ / SH, 0.33333));
                if (Z1 < ZC && UN > 0.0) UN = 0.0;
                f.f0 = copysign(CE * K::C1 * pow_rare(SH, 1.5), (double)(Z1 - ZC));
                f.f1 = f.f0 * fabs(UN);
                f.f2 = f.f0 * VT;
            } else {
                R HD = ZC - TOP;
                R UN = fmax(QL1, (R)0.0);
                R VT = QL2;
                R SH = HD + DZ;
                R CE = fmin(1.0, 1.05 * pow_rare(DZ / SH, 0.33333));
                f.f0 = copysign(CE * K::C1 * pow_rare(SH, 1.5), (double)(Z1 - ZC));
                f.f1 = (Z1 - ZC) * fabs(UN) * UN;
                f.f2 = (Z1 - ZC) * fabs(UN) * VT;
            }
            f.f3 = 4.905 * pow_rare(TOP - ZBp, 2.0);
        }
    } break;
    case K_B7: {
        if (!RARE) break;
        const R TOP = (R)bc.BF[bc.cell];
        const R HB = 0;   // the reference reads an uninitialised HB here
        if (Z1 > TOP || ZC > TOP) {
            *klas7_fired = true;
            R CQ = qd<R>(Z1, ZC, TOP);
            R CB = K::BRDTH / bc.SIDE;
            f.f0 = CQ * CB;
            f.f1 = CB * copysign(CQ * CQ / HB, CQ);
            f.f3 = 4.905 * HB * HB;
        } else {
            f.f0 = 0; f.f1 = 0; f.f2 = 0;
            f.f3 = 4.905 * H1 * H1;
        }
    } break;
    default:  // K_BUNK: FLUX0/1/3 keep their previous values
        break;
    }
    return f;
}

// ---------------------------------------------------------------------------
// Wet/dry ladder of an interior slot (F:113-147), seen from the own cell.
// Returns the branch: 0 both dry, 1 ZI<=BC, 2 ZC<=ZBp, 3/4 H1<=HM2 (inflow /
// outflow), 5/6 HC<=HM2, 7 both wet (Riemann problem, F:149-200).  For
// branches 0..6 *f receives the flux.  At most one pow(., 1.5) and one sqrt
// per slot, so their expansions appear once in the code.
// The branch alone (the task planner groups slots by it).
template <class R>
__host__ __device__ __forceinline__ int ladder_branch(const Scalars<R> &S, R H1, R ZI, R ZBp, R HC, R BC, R ZC) {
    int br;
    if (H1 <= S.HM1 && HC <= S.HM1) br = 0;
    else if (ZI <= BC) br = 1;
    else if (ZC <= ZBp) br = 2;
    else if (H1 <= S.HM2) br = (ZC > ZI) ? 3 : 4;
    else if (HC <= S.HM2) br = (ZI > ZC) ? 5 : 6;
    else br = 7;
    return br;
}

template <class R>
__host__ __device__ __forceinline__ int ladder(const Scalars<R> &S, R H1, R ZI, R ZBp, R HC, R BC, R ZC, R UC,
                                               R VC, R QL1, R QL2, R COSJ, R SINJ, Flux<R> *f) {
    using K = Consts<R>;
    const int br = ladder_branch<R>(S, H1, ZI, ZBp, HC, BC, ZC);
    if (br == 7) return br;
    // Branch sets as bit masks over br (no multi-case switch: see osher):
    // pow(., 1.5) for {1, 2, 4, 6} of HC in {1, 6}, sqrt for {3, 5}.
    const bool use_p = (0x56 >> br) & 1, p_hc = (0x42 >> br) & 1, use_sq = (0x28 >> br) & 1;
    // pow in double, kept in double until it meets C1 (F:119-145 multiply
    // the double result; rounding it to R first would round twice in fp32)
    using P = decltype(pow_rare(H1, 1.5));
    P p15 = 0;
    if (use_p) p15 = pow_rare(p_hc ? HC : H1, 1.5);
    R DH = 0, UN = 0;
    if (use_sq) {
        DH = fmax(br == 3 ? ZC - ZBp : ZI - BC, S.HM1);
        const R sq = sqrt_r<R>(DH);
        UN = (br == 3) ? -K::C1 * sq : K::C1 * sq;
    }
    // every branch's expressions, in the reference's evaluation order,
    // then selected (F:121-147)
    const R f0_dh = DH * UN;                    // 3, 5
    const R f0_pn = -K::C1 * p15;               // 1, 6
    const R f0_pp = K::C1 * p15;                // 2, 4
    const R f1_1 = H1 * QL1 * fabs(QL1);        // 1
    const R f1_2 = H1 * fabs(QL1) * QL1;        // 2
    const R f1_dh = f0_dh * UN;                 // 3, 5
    const R f1_6 = H1 * QL1 * QL1;              // 6
    const R f2_2 = H1 * fabs(QL1) * QL2;        // 2
    const R f2_3 = f0_dh * (VC * COSJ - UC * SINJ);
    const R f2_5 = f0_dh * QL2;
    const R HC1 = ZC - ZBp;
    const R f3_h = 4.905 * H1 * H1;             // 1, 3, 4, 6
    const R f3_5 = 4.905 * HC1 * HC1;           // 5
    f->f0 = br == 0 ? (R)0 : use_sq ? f0_dh : p_hc ? f0_pn : f0_pp;
    f->f1 = br == 1 ? f1_1 : br == 2 ? f1_2 : use_sq ? f1_dh : br == 6 ? f1_6 : (R)0;
    f->f2 = br == 2 ? f2_2 : br == 3 ? f2_3 : br == 5 ? f2_5 : (R)0;
    f->f3 = (br == 0 || br == 2) ? (R)0 : br == 5 ? f3_5 : f3_h;
    return br;
}

// Riemann problem of an interior edge in its primary orientation: left
// state a (the primary cell), right state b, normal (c, s) of a's slot
// (F:151-164).  Both sides of the edge use the same FO and f1 (F:165-200
// solves the neighbour's primary problem on the mirror side).
template <class R> struct EdgeSol {
    R FO0, FO2, FO3;
    R f1;    // FO[1] plus the depth-ratio correction
    R ZBa;   // bed of the primary cell (mirror side's F3)
};

template <class R>
__host__ __device__ __forceinline__ EdgeSol<R> edge_solve(const Scalars<R> &S, R Ha, R Ua, R Va, R ZBa, R Hb, R Ub,
                                                          R Vb, R Zb, R ZBb, R c, R s) {
    using K = Consts<R>;
    const R qa1 = Ua * c + Va * s;
    const R qa2 = Va * c - Ua * s;
    const R CLa = sqrt_r<R>(9.81 * Ha);
    const R FILa = qa1 + 2 * CLa;
    const R ZCb = fmax(ZBb, Zb);
    const R HCb = fmax(Hb, S.HM1);
    const R QR0 = fmax(ZCb - ZBa, S.HM1);
    const R UR = Ub * c + Vb * s;
    const R ratio = fmin(HCb / QR0, (R)1.5);
    R QR1 = UR * ratio;
    if (HCb <= S.HM2 || QR0 <= S.HM2) QR1 = copysign(K::VMIN, UR);
    const R QR2 = Vb * c - Ub * s;
    R FO[4];
    osher<R>(Ha, Ha, qa1, qa2, QR0, QR1, QR2, FILa, FO);
    EdgeSol<R> e;
    e.FO0 = FO[0];
    e.FO2 = FO[2];
    e.FO3 = FO[3];
    e.f1 = FO[1] + (1 - fmin(HCb / QR0, (R)1.5)) * HCb * UR * UR / 2;
    e.ZBa = ZBa;
    return e;
}

// Flux of the primary (prim) or mirror side of a solved edge; ZBp is the
// own cell's bed.
template <class R>
__host__ __device__ __forceinline__ Flux<R> edge_side(const EdgeSol<R> &e, bool prim, R ZBp) {
    Flux<R> f;
    f.f1 = e.f1;
    f.f2 = e.FO2;
    if (prim) {
        f.f0 = e.FO0;
        f.f3 = e.FO3;
    } else {
        f.f0 = -e.FO0;
        R ZA = sqrt_r<R>(e.FO3 / 4.905) + e.ZBa;
        R HC3 = 0;
        if (ZA > ZBp) HC3 = ZA - ZBp;
        f.f3 = 4.905 * HC3 * HC3;
    }
    return f;
}

// Flux of one non-padding slot (CalculateFluxKernel body F:47-202).
//   own:  H1 U1 V1 Z1, ZB1 (own ZB1 array value), ZBp (own bed)
//   nbr:  Hn Un Vn Zn ZBn, valid when has_nbr (NAC != 0)
template <class R, int VAR, bool RARE = true>
__host__ __device__ inline Flux<R> slot_flux(int kind, const Scalars<R> &S, const BcCtx<R> &bc,
                                             R H1, R U1, R V1, R Z1, R ZB1, R ZBp,
                                             bool has_nbr, R Hn, R Un, R Vn, R Zn, R ZBn,
                                             R COSJ, R SINJ, Flux<R> prev, bool *klas7_fired) {
    const R QL1 = U1 * COSJ + V1 * SINJ;
    const R QL2 = V1 * COSJ - U1 * SINJ;
    const R ZI = fmax(Z1, ZB1);
    R HC = 0, BC = 0, ZC = 0, UC = 0, VC = 0;
    if (has_nbr) {
        HC = fmax(Hn, S.HM1);
        BC = ZBn;
        ZC = fmax(BC, Zn);
        UC = Un;
        VC = Vn;
    }
    if (kind >= K_B10) {
        const R CL = sqrt_r<R>(9.81 * H1);
        const R FIL = QL1 + 2 * CL;
        return boundary_flux<R, VAR, RARE>(kind, S, bc, H1, U1, V1, Z1, ZBp, ZC, UC, VC, QL1, QL2, FIL,
                                           prev, klas7_fired);
    }
    Flux<R> f;
    if (ladder<R>(S, H1, ZI, ZBp, HC, BC, ZC, UC, VC, QL1, QL2, COSJ, SINJ, &f) < 7) return f;
    // Both sides wet: one Riemann solve on (left a, right b).  Primary side:
    // a = own, b = neighbour, normal (COSJ, SINJ).  Mirror side: a = neighbour,
    // b = own, normal negated -- the neighbour's own primary problem.
    const bool prim = (kind == K_PRIMARY);
    const R c = prim ? COSJ : -COSJ;
    const R s = prim ? SINJ : -SINJ;
    const EdgeSol<R> e = edge_solve<R>(S, prim ? H1 : Hn, prim ? U1 : Un, prim ? V1 : Vn, prim ? ZBp : ZBn,
                                       prim ? Hn : H1, prim ? Un : U1, prim ? Vn : V1, prim ? Zn : Z1,
                                       prim ? ZBn : ZBp, c, s);
    return edge_side<R>(e, prim, ZBp);
}

// Both slots of an interior edge whose sides are the primary cell P and the
// mirror cell M (normals exactly opposite), each through its own ladder and
// one shared Riemann solve: the same values as slot_flux on each side.
template <class R>
__host__ __device__ __forceinline__ void pair_flux(const Scalars<R> &S, R HP, R UP, R VP, R ZP, R ZB1P, R ZBP,
                                                   R HM, R UM, R VM, R ZM, R ZB1M, R ZBM, R c, R s,
                                                   Flux<R> *fP, Flux<R> *fM) {
    const R cm = -c, sm = -s;
    const R QL1P = UP * c + VP * s, QL2P = VP * c - UP * s;
    const R QL1M = UM * cm + VM * sm, QL2M = VM * cm - UM * sm;
    const int bP = ladder<R>(S, HP, fmax(ZP, ZB1P), ZBP, fmax(HM, S.HM1), ZBM, fmax(ZBM, ZM), UM, VM, QL1P, QL2P,
                             c, s, fP);
    const int bM = ladder<R>(S, HM, fmax(ZM, ZB1M), ZBM, fmax(HP, S.HM1), ZBP, fmax(ZBP, ZP), UP, VP, QL1M, QL2M,
                             cm, sm, fM);
    if (bP == 7 || bM == 7) {
        const EdgeSol<R> e = edge_solve<R>(S, HP, UP, VP, ZBP, HM, UM, VM, ZM, ZBM, c, s);
        if (bP == 7) *fP = edge_side<R>(e, true, ZBP);
        if (bM == 7) *fM = edge_side<R>(e, false, ZBM);
    }
}

// Sign-exact stand-in for the flux of a triangle's padding slot (NAC 0,
// SIDE COSF SINF 0, KLAS 0).  The reference computes a flux there and adds
// SIDE*FLUX = +-0 into the cell sums; only the signs of those zeros can
// matter (they decide the sign of an exactly-zero WU, which copysign later
// reads).  This evaluates the same branch ladder with every positive
// magnitude (pow, sqrt) replaced by a positive stand-in, so SIDE*f gives the
// reference's zero, bit for bit, without the transcendental work.
template <class R>
__host__ __device__ __forceinline__ Flux<R> padding_flux_signs(const Scalars<R> &S, R H1, R U1, R V1,
                                                               R Z1, R ZB1, R ZBp) {
    const R QL1 = U1 * (R)0 + V1 * (R)0;
    const R QL2 = V1 * (R)0 - U1 * (R)0;
    const R ZI = fmax(Z1, ZB1);
    Flux<R> f;
    if (H1 <= S.HM1) {                      // both dry (HC = 0)
        f.f0 = 0; f.f1 = 0; f.f2 = 0; f.f3 = 0;
    } else if (ZI <= (R)0) {                // BC = 0
        f.f0 = -(R)0;                       // -C1 * pow(0, 1.5)
        f.f1 = H1 * QL1 * fabs(QL1);
        f.f2 = 0;
        f.f3 = 1;
    } else if ((R)0 <= ZBp) {               // ZC = 0
        f.f0 = 1;
        f.f1 = H1 * fabs(QL1) * QL1;
        f.f2 = H1 * fabs(QL1) * QL2;
        f.f3 = 0;
    } else if (H1 <= S.HM2) {               // ZC = 0 < ZI here: no inflow form
        f.f0 = 1; f.f1 = 0; f.f2 = 0; f.f3 = 1;
    } else {                                // HC = 0 <= HM2 and ZI > ZC = 0
        f.f0 = 1; f.f1 = 1;
        f.f2 = QL2;                         // F0 * QL2 with F0 > 0
        f.f3 = 1;
    }
    return f;
}

// ---------------------------------------------------------------------------
// UpdateCellKernel arithmetic after the flux sums (F:744-783; variant B's
// per-cell CFL Fv0:751-768 and 5 m/s cap Fv0:791-793).
template <class R> struct CellOut { R H, U, V, Z, W; };

// W of UpdateCellKernel (F:783): a function of the new U and V only, so the
// fused paths recompute it when the state leaves the internal layout.
template <class R> __host__ __device__ __forceinline__ R speed(R U, R V) { return sqrt_r<R>(U * U + V * V); }

// DT / AREA of the cell: the fixed step (variant A) or the per-cell CFL
// step of variant B (Fv0:751-768).
template <class R, int VAR>
__host__ __device__ __forceinline__ R cell_dta(const Scalars<R> &S, R H1, R U1, R AREA, R s0, R s1, R s2, R s3) {
    if (VAR == 1) {
        R SIDEX;
        if (s3 > (R)0.0) {
            SIDEX = fmin((R)0.5 * (s0 + s2), (R)0.5 * (s1 + s3));
        } else {
            R SIDES = (R)0.5 * (s0 + s1 + s2);
            SIDEX = sqrt_r<R>((SIDES - s0) * (SIDES - s1) * (SIDES - s2) / SIDES);
        }
        R HSIDE = fmax(H1, S.HM1);
        R DT2 = SIDEX / (U1 + sqrt_r<R>((R)9.81 * HSIDE));
        DT2 = fmin((R)S.DT, DT2);
        DT2 = fmax(DT2, (R)S.DT / (R)10.0);
        return DT2 / AREA;
    }
    return 1.0 * S.DT / (1.0 * AREA);
}

// The update given DT / AREA (F:744-783).
template <class R, int VAR>
__host__ __device__ __forceinline__ CellOut<R> update_with_dta(const Scalars<R> &S, R H1, R U1, R V1, R ZBp, R DTA,
                                                               R FNC, R WH, R WU, R WV) {
    using K = Consts<R>;
    const R WDTA = 1.00 * DTA;
    CellOut<R> o;
    o.H = fmax(H1 - WDTA * WH + K::QLUA, S.HM1);
    o.Z = o.H + ZBp;
    if (o.H <= S.HM1) {
        o.U = 0.0;
        o.V = 0.0;
    } else if (o.H <= S.HM2) {
        o.U = copysign(fmin(K::VMIN, fabs(U1)), U1);
        o.V = copysign(fmin(K::VMIN, fabs(V1)), V1);
    } else {
        const R QX1 = H1 * U1, QY1 = H1 * V1;
        const R DTAU = WDTA * WU, DTAV = WDTA * WV;
        const R WSF = FNC * sqrt_r<R>(U1 * U1 + V1 * V1) / pow_d(H1, 0.33333);
        R U2 = (QX1 - DTAU - S.DT * WSF * U1) / o.H;
        R V2 = (QY1 - DTAV - S.DT * WSF * V1) / o.H;
        const R cap = (VAR == 1) ? (R)5.0 : (R)15.0;
        o.U = copysign(fmin(fabs(U2), cap), U2);
        o.V = copysign(fmin(fabs(V2), cap), V2);
    }
    o.W = speed<R>(o.U, o.V);
    return o;
}

template <class R, int VAR>
__host__ __device__ __forceinline__ CellOut<R> update_math(const Scalars<R> &S, R H1, R U1, R V1,
                                                           R ZBp, R AREA, R FNC, R WH, R WU, R WV,
                                                           R s0, R s1, R s2, R s3) {
    return update_with_dta<R, VAR>(S, H1, U1, V1, ZBp, cell_dta<R, VAR>(S, H1, U1, AREA, s0, s1, s2, s3), FNC,
                                   WH, WU, WV);
}

// One slot's terms of the cell sums WH, WU, WV (F:732-742).
template <class R> struct Contrib { R h, u, v; };
template <class R>
__host__ __device__ __forceinline__ Contrib<R> contrib(const Flux<R> &f, R SL, R SLCA, R SLSA) {
    const R FLR_1 = f.f1 + f.f3;
    const R FLR_2 = f.f2;
    return Contrib<R>{SL * f.f0, SLCA * FLR_1 - SLSA * FLR_2, SLSA * FLR_1 + SLCA * FLR_2};
}

// Accumulate one slot into the cell sums in the reference's order.
template <class R>
__host__ __device__ __forceinline__ void accumulate(R &WH, R &WU, R &WV, const Flux<R> &f, R SL, R SLCA,
                                                    R SLSA) {
    const Contrib<R> c = contrib<R>(f, SL, SLCA, SLSA);
    WH += c.h;
    WU += c.u;
    WV += c.v;
}

}  // namespace hydro
