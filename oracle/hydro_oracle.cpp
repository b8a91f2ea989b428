// hydro_oracle.cpp -- CPU restatement of the hydro-cal shallow-water timestep.
//
// TEST INFRASTRUCTURE.  This file is the correctness oracle.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
// only as the checker (or as the timed CPU baseline); the product library
// never links or calls it.
//
// It follows the reference kernels slot by slot and statement by statement:
//   variant A  cuda_native_impl/hydro-cal-src/src/functors.cu            (F)
//   variant B  cuda_native_impl/baselines/native_cuda_v0/fp64/.../functors.cu (Fv0)
//   fp32       cuda_native_impl/fp32_src/hydro-cal-src/src/functors.cu  (F32)
// The reference compiles one source per precision with `using Real = double`
// or `float` (include/common.hpp:15 of each tree) and leaves the literals
// (4.905, 9.81, 39.24, 6.264, 313.92, 1.5, 0.33333, ...) as double, so in the
// fp32 build every sub-expression touching a literal is evaluated in double
// and rounded to float on assignment.  Writing the same expressions in a
// template over Real reproduces those promotions by the C++ rules themselves.
// The fp32 tree calls sqrtf() where the fp64 tree calls sqrt(): sqrt_r()
// below rounds its argument to Real first, which is what sqrtf(double) does.
//
// Build: see oracle/Makefile (g++ -O2 -ffp-contract=off; no fast-math).  The
// fp64 reference build lets nvcc contract a*b+c into FMAs; this oracle does
// not, which is why fp64 parity is stated to 1e-12 relative and not bitwise.
//
// Undefined behaviour of the reference, resolved as the Taichi port resolves
// it (taichi_impl/F2_hydro_taichi_fp64.py):
//   * OSHER K2 when UA == CA (or NaN): 4           (F:586-597 vs T2:180-187)
//   * OSHER K1 when no branch matches (NaN): 4     (F:599-610 vs T2:190-197)
//   * KLAS 7 reads an uninitialised HB: HB = 0     (F:525,562-565 vs T2:555)
//   * LAQP with no bracketing interval: Y = 0      (F:423-443 vs T2:393-412)
// Slots whose KLAS is a boundary code without a handler (2, 8, >= 11) write
// only FLUX2 and keep FLUX0/1/3 from before (F:524-566), as here.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>

// Planted mistakes for the mutation test of the oracle's pins
// (tests/test_oracle_mutants.py): a build with -DORACLE_MUTANT=k makes
// mistake k below, each one a plausible slip in restating the reference;
// the pins must catch every one.  Every other build has ORACLE_MUTANT 0 and
// none of them.
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif
#define MUTANT(k) (ORACLE_MUTANT == (k))

extern "C" {

// Host-side mesh view: same fields and meaning as hydro_mesh_desc in
// include/hydro.h, host pointers.
struct oracle_mesh {
    int32_t dtype;    // 0 = double, 1 = float
    int32_t variant;  // 0 = A, 1 = B
    int32_t CELL, NDAYS, NHQ, MDT;
    double DT, HM1, HM2, JL;
    void *H, *U, *V, *Z, *W;
    const void *ZBC, *ZB1, *AREA, *FNC, *BoundaryFeature;
    const int32_t *NHQ1;
    const void *ZW, *QW, *QT, *DQT, *ZT, *DZT;
    const int32_t *NAC;
    int32_t *KLAS;
    const void *SIDE, *COSF, *SINF, *SLCOS, *SLSIN;
    void *FLUX0, *FLUX1, *FLUX2, *FLUX3;
};

}  // extern "C"

namespace {

template <class R>
struct View {
    int CELL, NDAYS, NHQ, variant;
    R DT, HM1, HM2, JL;
    R *H, *U, *V, *Z, *W;
    const R *ZBC, *ZB1, *AREA, *FNC, *BF;
    const int32_t *NHQ1;
    const R *ZW, *QW, *QT, *DQT, *ZT, *DZT;
    const int32_t *NAC;
    int32_t *KLAS;
    const R *SIDE, *COSF, *SINF, *SLCOS, *SLSIN;
    R *F0, *F1, *F2, *F3;

    explicit View(const oracle_mesh &m)
        : CELL(m.CELL), NDAYS(m.NDAYS), NHQ(m.NHQ), variant(m.variant),
          DT((R)m.DT), HM1((R)m.HM1), HM2((R)m.HM2), JL((R)m.JL),
          H((R *)m.H), U((R *)m.U), V((R *)m.V), Z((R *)m.Z), W((R *)m.W),
          ZBC((const R *)m.ZBC), ZB1((const R *)m.ZB1), AREA((const R *)m.AREA),
          FNC((const R *)m.FNC), BF((const R *)m.BoundaryFeature), NHQ1(m.NHQ1),
          ZW((const R *)m.ZW), QW((const R *)m.QW), QT((const R *)m.QT),
          DQT((const R *)m.DQT), ZT((const R *)m.ZT), DZT((const R *)m.DZT),
          NAC(m.NAC), KLAS(m.KLAS), SIDE((const R *)m.SIDE), COSF((const R *)m.COSF),
          SINF((const R *)m.SINF), SLCOS((const R *)m.SLCOS), SLSIN((const R *)m.SLSIN),
          F0((R *)m.FLUX0), F1((R *)m.FLUX1), F2((R *)m.FLUX2), F3((R *)m.FLUX3) {}
};

// modelConstants (include/common.hpp:44-52): typed Real in the reference.
template <class R> struct Kc {
    static constexpr R C0 = (R)1.33;
    static constexpr R C1 = (R)1.7;
    static constexpr R VMIN = (R)0.001;
    static constexpr R QLUA = (R)0.0;
    static constexpr R BRDTH = (R)100.0;
};

// sqrt in the fp64 tree, sqrtf in the fp32 tree (argument rounded to Real).
template <class R> inline R sqrt_r(double x) { return std::sqrt((R)x); }
template <> inline double sqrt_r<double>(double x) { return std::sqrt(x); }

// QF (F:505-511): physical flux of state (h, u, v) across a unit normal.
template <class R> inline void qf(R h, R u, R v, R F[4]) {
    F[0] = h * u;
    F[1] = F[0] * u;
    F[2] = F[0] * v;
    F[3] = 4.905 * h * h;
}

// Osher 16-case table (F:612-711): for each (K1, K2) the QS kinds applied in
// order, with sign; 0 terminates.  Kinds: 1 left state, 2 sonic point of the
// left curve, 3 intermediate state (left side), 5 intermediate state (right
// side), 6 sonic point of the right curve, 7 right state.
const signed char kOsherSeq[4][4][6] = {
    /* K1=1 */ {{2, 0}, {3, 0}, {5, 0}, {6, 0}},
    /* K1=2 */ {{1, 0}, {1, -2, 3, 0}, {1, -2, 5, 0}, {1, -2, 6, 0}},
    /* K1=3 */ {{2, -6, 7, 0}, {3, -6, 7, 0}, {5, -6, 7, 0}, {7, 0}},
    /* K1=4 */ {{1, -6, 7, 0}, {1, -2, 3, -6, 7, 0}, {1, -2, 5, -6, 7, 0}, {1, -2, 7, 0}},
};

// QS<T> (F:445-503).  Kinds 3 and 5 update FIL / FIR in place; later kinds
// of the same sequence see the updated value (F:467, F:479; also MF:1288,1297).
template <class R>
inline void qs(int kind, int j, const R QL[3], const R QR[3], R &FIL, R &FIR, R acc[4]) {
    R F[4];
    switch (kind) {
    case 1: qf<R>(QL[0], QL[1], QL[2], F); break;
    case 2: { R US = FIL / 3; R HS = US * US / 9.81; qf<R>(HS, US, QL[2], F); } break;
    case 3: { R UA = (FIL + FIR) / 2; FIL = FIL - UA; R HA = FIL * FIL / 39.24; qf<R>(HA, UA, QL[2], F); } break;
    case 5: { R UA = (FIL + FIR) / 2; FIR = FIR - UA; R HA = FIR * FIR / 39.24; qf<R>(HA, UA, QR[2], F); } break;
    case 6: { R US = FIR / 3; R HS = US * US / 9.81; qf<R>(HS, US, QR[2], F); } break;
    case 7: qf<R>(QR[0], QR[1], QR[2], F); break;
    default: return;
    }
    for (int i = 0; i < 4; i++) acc[i] += F[i] * j;
}

// OSHER (F:570-712).
template <class R>
void osher(R H_pre, const R QL[3], const R QR[3], R FIL, R out[4]) {
    R CR = sqrt_r<R>(9.81 * QR[0]);
    R FIR = QR[1] - 2 * CR;
    R UA = (FIL + FIR) / 2;
    R CA = std::fabs((FIL - FIR) / 4);
    R CL = sqrt_r<R>(9.81 * H_pre);
    out[0] = out[1] = out[2] = out[3] = 0;
    int K2;
    if (CA < UA) K2 = 1;
    else if (UA >= 0.0 && UA < CA) K2 = 2;
    else if (UA >= -CA && UA < 0.0) K2 = 3;
    else K2 = 4;
    int K1;
    if (QL[1] < CL && QR[1] >= -CR) K1 = 1;
    else if (QL[1] >= CL && QR[1] >= -CR) K1 = 2;
    else if (QL[1] < CL && QR[1] < -CR) K1 = 3;
    else K1 = 4;
    const signed char *seq = kOsherSeq[K1 - 1][K2 - 1];
    for (int s = 0; s < 6 && seq[s] != 0; s++) {
        int kind = seq[s] < 0 ? -seq[s] : seq[s];
        int sign = seq[s] < 0 ? -1 : 1;
        qs<R>(kind, sign, QL, QR, FIL, FIR, out);
    }
}

// QD (F:394-421): submerged / free weir discharge.
template <class R> R qd(R ZL, R ZR, R ZB) {
    const R CM = 0.384, SIGMA = 0.667, FI = 4.43;
    R ZU = std::fmax(ZL, ZR), ZD = std::fmin(ZL, ZR);
    R H0 = ZU - ZB, HS = ZD - ZB;
    R DELTA = HS / H0;
    R out;
    if (DELTA <= SIGMA) {
        out = std::copysign(CM * std::pow((double)(MUTANT(9) ? HS : H0), 1.5), (double)(ZL - ZR));
    } else {
        R DH = ZU - ZD;
        if (DH > 0.09) out = std::copysign(FI * HS * sqrt_r<R>(DH), ZL - ZR);
        else out = std::copysign(FI * HS * 0.3 * DH / 0.1, (double)(ZL - ZR));
    }
    return out;
}

// LAQP (F:423-443): piecewise-linear lookup of Y(X) on the table (A, B).
template <class R> R laqp(R X, const R *A, const R *B, int MS) {
    R Y = 0;
    if (MS <= 0) return Y;
    if (X < A[0]) return B[0];
    if (X > A[MS - 1]) return B[MS - 1];
    for (int i = 0; i < MS - 1; i++)
        if (X >= A[i] && X <= A[i + 1])
            return MUTANT(10) ? B[i] : B[i] + (B[i + 1] - B[i]) / (A[i + 1] - A[i]) * (X - A[i]);
    return Y;
}

// Iterated boundary Riemann invariant shared by KLAS 1 and KLAS 3
// (F:355-364, F:379-388): returns URB.
template <class R> R urb_iterate(R QLn, R H_pre, R HB1, double tol) {
    R FIAL = QLn + 6.264 * sqrt_r<R>(H_pre);
    R UR0 = QLn;
    R URB = UR0;
    for (int it = 1; it <= 30; it++) {
        R FIAR = URB - 6.264 * sqrt_r<R>(HB1);
        URB = (FIAL + FIAR) * (FIAL - FIAR) * (FIAL - FIAR) / HB1 / 313.92;
        if (std::fabs(URB - UR0) <= tol) break;
        UR0 = URB;
    }
    return URB;
}

// BOUNDA (F:513-568) and the KLAS handlers (F:204-392).
template <class R>
void bounda(View<R> &m, int kt, int jt, int pos, int idx, R H_pre, R U_pre, R V_pre, R Z_pre,
            R ZC, R UC, R VC, R FIL, R QL[3]) {
    using K = Kc<R>;
    R CL = sqrt_r<R>(9.81 * H_pre);
    bool super = (m.variant == 1) ? (QL[1] > CL) : (QL[1] > CL && H_pre < m.HM2);
    if (super) {
        m.F0[idx] = H_pre * QL[1];
        m.F1[idx] = m.F0[idx] * QL[1];
        m.F2[idx] = m.F0[idx] * QL[2];
        m.F3[idx] = 4.905 * H_pre * H_pre;
        return;
    }
    m.F2[idx] = 0;
    if (QL[1] > 0) m.F2[idx] = H_pre * QL[1] * QL[2];
    const int kl = m.KLAS[idx];
    const size_t row = (size_t)jt * (size_t)m.CELL + (size_t)pos;
    if (kl == 10) {
        m.F0[idx] = -(m.QT[row] + m.DQT[row] * kt);
        m.F0[idx] = m.F0[idx] / m.SIDE[idx];
        R QB2 = m.F0[idx] * m.F0[idx];
        R HB0 = H_pre;
        R HB = 0;
        for (int k = 1; k <= 20; k++) {
            R Wt = FIL - m.F0[idx] / HB0;
            HB = Wt * Wt / 39.24;
            if (std::fabs(HB0 - HB) <= 0.005) break;
            HB0 = MUTANT(5) ? HB0 : HB0 * 0.5 + HB * 0.5;
        }
        m.F1[idx] = (HB <= 1) ? (R)0 : QB2 / (MUTANT(4) ? H_pre : HB);
        m.F2[idx] = 0;
        m.F3[idx] = 4.905 * HB * HB;
    } else if (kl == 3) {
        R QZH3 = (U_pre * m.COSF[idx] + V_pre * m.SINF[idx]) * H_pre * m.SIDE[idx];
        QZH3 = std::fmax(QZH3, (R)0.0);
        const R *QWr = m.QW + (size_t)pos * (size_t)m.NHQ;
        const R *ZWr = m.ZW + (size_t)pos * (size_t)m.NHQ;
        R ZQH1 = laqp<R>(QZH3, QWr, ZWr, m.NHQ1[pos]);
        R ZQH11 = ZQH1 - Z_pre;
        R HB1 = ZQH1 - m.ZBC[pos];
        HB1 = std::fmax(HB1, m.HM2);
        QL[1] = std::fmax(QL[1], (R)0.0);
        if (QZH3 <= QWr[1] || HB1 <= m.HM2 || ZQH11 >= 0.1) {
            m.F0[idx] = H_pre * QL[1];
            m.F1[idx] = m.F0[idx] * QL[1];
            m.F3[idx] = 4.905 * H_pre * H_pre;
        } else {
            R URB = urb_iterate<R>(QL[1], H_pre, HB1, 0.001);
            m.F0[idx] = HB1 * URB;
            m.F1[idx] = m.F0[idx] * URB;
            m.F3[idx] = 4.905 * HB1 * HB1;
        }
    } else if (kl == 1) {
        R HB1 = m.ZT[row] + m.DZT[row] * (m.variant == 1 ? kt : jt) - m.ZBC[pos];
        HB1 = std::fmax(HB1, m.HM2);
        R URB = urb_iterate<R>(MUTANT(6) ? QL[2] : QL[1], H_pre, HB1, 0.0001);
        m.F0[idx] = HB1 * URB;
        m.F1[idx] = m.F0[idx] * URB;
        m.F3[idx] = 4.905 * HB1 * HB1;
    } else if (kl == 4) {
        m.F0[idx] = 0;
        m.F1[idx] = 0;
        m.F2[idx] = 0;
        m.F3[idx] = 4.905 * H_pre * H_pre;
    } else if (kl == 5) {
        QL[1] = std::fmax(QL[1], (R)0.0);
        m.F0[idx] = H_pre * QL[1];
        m.F1[idx] = m.F0[idx] * QL[1];
        m.F3[idx] = 4.905 * H_pre * H_pre * (1.0 - m.JL) * (MUTANT(7) ? 1.0 : (1.0 - m.JL));
    } else if (kl == 6) {
        R TOP = m.BF[pos];
        R ZB = m.ZBC[pos];
        R cs = m.COSF[idx], sn = m.SINF[idx];
        if (Z_pre <= TOP && ZC <= TOP) {
            m.F0[idx] = 0; m.F1[idx] = 0; m.F2[idx] = 0;
            m.F3[idx] = 4.905 * H_pre * H_pre;
        } else if (Z_pre > TOP && ZC <= TOP) {
            m.F0[idx] = (MUTANT(8) ? K::C1 : K::C0) * std::pow((double)(Z_pre - TOP), 1.5);
            m.F1[idx] = m.F0[idx] * QL[1];
            m.F2[idx] = m.F0[idx] * QL[2];
            m.F3[idx] = 4.905 * std::pow((double)(TOP - ZB), 2);
        } else if (Z_pre <= TOP && ZC > TOP) {
            m.F0[idx] = -K::C0 * std::pow((double)(ZC - TOP), 1.5);
            m.F1[idx] = m.F0[idx] * std::fmin(UC * cs + VC * sn, (R)0.0);
            m.F2[idx] = m.F0[idx] * (VC * cs - UC * sn);
            m.F3[idx] = 4.905 * std::pow((double)(Z_pre - ZB), 2);
        } else {
            R DZ = std::fabs(Z_pre - ZC);
            if (Z_pre <= ZC) {
                R HD = Z_pre - TOP;
                R UN = std::fmin(UC * cs + VC * sn, (R)0.0);
                R VT = VC * cs - UC * sn;
                R SH = HD + DZ;
                R CE = std::fmin(1.0, 1.05 * std::pow((double)(DZ / SH), 0.33333));
                if (Z_pre < ZC && UN > 0.0) UN = 0.0;
                m.F0[idx] = std::copysign(CE * K::C1 * std::pow((double)SH, 1.5), (double)(Z_pre - ZC));
                m.F1[idx] = m.F0[idx] * std::fabs(UN);
                m.F2[idx] = m.F0[idx] * VT;
                m.F3[idx] = 4.905 * std::pow((double)(TOP - ZB), 2);
            } else {
                R HD = ZC - TOP;
                R UN = std::fmax(QL[1], (R)0.0);
                R VT = QL[2];
                R SH = HD + DZ;
                R CE = std::fmin(1.0, 1.05 * std::pow((double)(DZ / SH), 0.33333));
                m.F0[idx] = std::copysign(CE * K::C1 * std::pow((double)SH, 1.5), (double)(Z_pre - ZC));
                m.F1[idx] = (Z_pre - ZC) * std::fabs(UN) * UN;
                m.F2[idx] = (Z_pre - ZC) * std::fabs(UN) * VT;
                m.F3[idx] = 4.905 * std::pow((double)(TOP - ZB), 2);
            }
        }
    } else if (kl == 7) {
        R TOP = m.BF[pos];
        R HB = 0;
        if (Z_pre > TOP || ZC > TOP) {
            if (!MUTANT(11)) m.KLAS[idx] = 0;
            R CQ = qd<R>(Z_pre, ZC, TOP);
            R CB = K::BRDTH / m.SIDE[idx];
            m.F0[idx] = CQ * CB;
            m.F1[idx] = CB * std::copysign(CQ * CQ / HB, CQ);
            m.F3[idx] = 4.905 * HB * HB;
        } else {
            m.F0[idx] = 0; m.F1[idx] = 0; m.F2[idx] = 0;
            m.F3[idx] = 4.905 * H_pre * H_pre;
        }
    }
    // any other boundary code: only FLUX2 was written
}

// CalculateFluxKernel body for one slot (F:47-202).
template <class R> void flux_slot(View<R> &m, int kt, int jt, int idx) {
    using K = Kc<R>;
    const int pos = idx / 4;
    R U1 = m.U[pos], V1 = m.V[pos], H1 = m.H[pos], Z1 = m.Z[pos], ZB1 = m.ZB1[pos];
    int NC = m.NAC[idx] - 1;
    int KP = m.KLAS[idx];
    R COSJ = m.COSF[idx], SINJ = m.SINF[idx];
    R QL[3];
    QL[0] = H1;
    QL[1] = U1 * COSJ + V1 * SINJ;
    QL[2] = V1 * COSJ - U1 * SINJ;
    R CL = sqrt_r<R>(9.81 * H1);
    R FIL = QL[1] + 2 * CL;
    R HC, BC, ZC, UC, VC;
    R ZI = std::fmax(Z1, ZB1);
    if (NC == -1) {
        HC = 0; BC = 0; ZC = 0; UC = 0; VC = 0;
    } else {
        HC = std::fmax(m.H[NC], m.HM1);
        BC = m.ZBC[NC];
        ZC = std::fmax(BC, m.Z[NC]);
        UC = m.U[NC];
        VC = m.V[NC];
    }
    R *F0 = m.F0, *F1 = m.F1, *F2 = m.F2, *F3 = m.F3;
    const R ZBp = m.ZBC[pos];
    if ((KP >= 1 && KP <= 8) || KP >= 10) {
        bounda<R>(m, kt, jt, pos, idx, H1, U1, V1, Z1, ZC, UC, VC, FIL, QL);
    } else if (H1 <= m.HM1 && HC <= m.HM1) {
        F0[idx] = 0; F1[idx] = 0; F2[idx] = 0; F3[idx] = 0;
    } else if (ZI <= BC) {
        F0[idx] = -K::C1 * std::pow((double)(MUTANT(12) ? H1 : HC), 1.5);
        F1[idx] = H1 * QL[1] * std::fabs(QL[1]);
        F2[idx] = 0;
        F3[idx] = 4.905 * H1 * H1;
    } else if (ZC <= ZBp) {
        F0[idx] = K::C1 * std::pow((double)H1, 1.5);
        F1[idx] = H1 * std::fabs(QL[1]) * QL[1];
        F2[idx] = H1 * std::fabs(QL[1]) * QL[2];
        F3[idx] = 0;
    } else if (MUTANT(3) && HC <= m.HM2) {   // the neighbour's shallow test first
        if (ZI > ZC) {
            R DH = std::fmax(ZI - BC, m.HM1);
            R UN = K::C1 * sqrt_r<R>(DH);
            R HC1 = ZC - ZBp;
            F0[idx] = DH * UN;
            F1[idx] = F0[idx] * UN;
            F2[idx] = F0[idx] * QL[2];
            F3[idx] = 4.905 * HC1 * HC1;
        } else {
            F0[idx] = -K::C1 * std::pow((double)HC, 1.5);
            F1[idx] = H1 * QL[1] * QL[1];
            F2[idx] = 0;
            F3[idx] = 4.905 * H1 * H1;
        }
    } else if (H1 <= m.HM2) {
        if (ZC > ZI) {
            R DH = std::fmax(ZC - ZBp, m.HM1);
            R UN = -K::C1 * sqrt_r<R>(DH);
            F0[idx] = DH * UN;
            F1[idx] = F0[idx] * UN;
            F2[idx] = F0[idx] * (VC * COSJ - UC * SINJ);
            F3[idx] = 4.905 * H1 * H1;
        } else {
            F0[idx] = K::C1 * std::pow((double)(MUTANT(1) ? HC : H1), 1.5);
            F1[idx] = 0; F2[idx] = 0;
            F3[idx] = 4.905 * H1 * H1;
        }
    } else if (HC <= m.HM2) {
        if (ZI > ZC) {
            R DH = std::fmax(ZI - BC, m.HM1);
            R UN = K::C1 * sqrt_r<R>(DH);
            R HC1 = ZC - (MUTANT(2) ? BC : ZBp);
            F0[idx] = DH * UN;
            F1[idx] = F0[idx] * UN;
            F2[idx] = F0[idx] * QL[2];
            F3[idx] = 4.905 * HC1 * HC1;
        } else {
            F0[idx] = -K::C1 * std::pow((double)HC, 1.5);
            F1[idx] = H1 * QL[1] * QL[1];
            F2[idx] = 0;
            F3[idx] = 4.905 * H1 * H1;
        }
    } else if (KP == 0 && pos < NC) {
        // primary side: solve this edge's Riemann problem directly (F:151-164)
        R QR[3], FO[4];
        QR[0] = std::fmax(ZC - ZBp, m.HM1);
        R UR = UC * COSJ + VC * SINJ;
        QR[1] = UR * std::fmin(HC / QR[0], (R)1.5);
        if (HC <= m.HM2 || QR[0] <= m.HM2) QR[1] = std::copysign(K::VMIN, UR);
        QR[2] = VC * COSJ - UC * SINJ;
        osher<R>(H1, QL, QR, FIL, FO);
        F0[idx] = FO[0];
        F1[idx] = FO[1] + (1 - std::fmin(HC / QR[0], (R)1.5)) * HC * UR * UR / 2;
        F2[idx] = FO[2];
        F3[idx] = FO[3];
    } else {
        // mirror side: re-solve the neighbour's problem in its frame (F:165-200)
        R HC2 = std::fmax(H1, m.HM1);
        R ZC1 = std::fmax(ZBp, Z1);
        R CL1 = sqrt_r<R>(9.81 * m.H[NC]);
        R COSJ1 = -COSJ, SINJ1 = -SINJ;
        R QL1[3], QR1[3], FO[4];
        QL1[0] = m.H[NC];
        QL1[1] = m.U[NC] * COSJ1 + m.V[NC] * SINJ1;
        QL1[2] = m.V[NC] * COSJ1 - m.U[NC] * SINJ1;
        R FIL1 = QL1[1] + 2 * CL1;
        QR1[0] = std::fmax(ZC1 - m.ZBC[NC], m.HM1);
        R UR1 = U1 * COSJ1 + V1 * SINJ1;
        QR1[1] = UR1 * std::fmin(HC2 / QR1[0], (R)1.5);
        if (HC2 <= m.HM2 || QR1[0] <= m.HM2) QR1[1] = std::copysign(K::VMIN, UR1);
        QR1[2] = V1 * COSJ1 - U1 * SINJ1;
        osher<R>(m.H[NC], QL1, QR1, FIL1, FO);
        F1[idx] = FO[1] + (1 - std::fmin(HC2 / QR1[0], (R)1.5)) * HC2 * UR1 * UR1 / 2;
        F0[idx] = -FO[0];
        F2[idx] = FO[2];
        R ZA = sqrt_r<R>(FO[3] / 4.905) + BC;
        R HC3 = 0;
        if (ZA > ZBp) HC3 = ZA - ZBp;
        F3[idx] = 4.905 * HC3 * HC3;
    }
}

// UpdateCellKernel body for one cell (F:717-784; variant B Fv0:748-802).
template <class R> void update_one(View<R> &m, int pos) {
    using K = Kc<R>;
    R WH = 0.0, WU = 0.0, WV = 0.0;
    for (int idx = 4 * pos; idx < 4 * (pos + 1); idx++) {
        R FLR_1 = m.F1[idx] + m.F3[idx];
        R FLR_2 = m.F2[idx];
        R SL = m.SIDE[idx], SLCA = m.SLCOS[idx], SLSA = m.SLSIN[idx];
        WH += SL * m.F0[idx];
        WU += SLCA * FLR_1 - SLSA * FLR_2;
        WV += SLSA * FLR_1 + SLCA * FLR_2;
    }
    R H1 = m.H[pos], U1 = m.U[pos], V1 = m.V[pos];
    R DTA;
    if (m.variant == 1) {
        // per-cell CFL step (Fv0:751-768; fp32 tree uses float literals here)
        R s0 = m.SIDE[4 * pos + 0], s1 = m.SIDE[4 * pos + 1];
        R s2 = m.SIDE[4 * pos + 2], s3 = m.SIDE[4 * pos + 3];
        R SIDEX;
        if (s3 > (R)0.0) {
            SIDEX = std::fmin((R)0.5 * (s0 + s2), (R)0.5 * (s1 + s3));
        } else {
            R SIDES = (R)0.5 * (s0 + s1 + s2);
            SIDEX = sqrt_r<R>((SIDES - s0) * (SIDES - s1) * (SIDES - s2) / SIDES);
        }
        R HSIDE = std::fmax(H1, m.HM1);
        R DT2 = SIDEX / (U1 + sqrt_r<R>((R)9.81 * HSIDE));
        DT2 = std::fmin((R)m.DT, DT2);
        DT2 = std::fmax(DT2, (R)m.DT / (R)10.0);
        DTA = DT2 / m.AREA[pos];
    } else {
        DTA = 1.0 * m.DT / (1.0 * m.AREA[pos]);
    }
    R WDTA = 1.00 * DTA;
    R H2, U2, V2, Z2, W2;
    H2 = std::fmax(H1 - WDTA * WH + K::QLUA, m.HM1);
    Z2 = H2 + m.ZBC[pos];
    if (H2 <= m.HM1) {
        U2 = 0.0;
        V2 = 0.0;
    } else if (H2 <= m.HM2) {
        U2 = std::copysign(std::fmin(K::VMIN, std::fabs(U1)), U1);
        V2 = std::copysign(std::fmin(K::VMIN, std::fabs(V1)), V1);
    } else {
        R QX1 = H1 * U1, QY1 = H1 * V1;
        R DTAU = WDTA * WU, DTAV = WDTA * WV;
        R WSF = m.FNC[pos] * sqrt_r<R>(U1 * U1 + V1 * V1) / std::pow((double)H1, 0.33333);
        U2 = (QX1 - DTAU - m.DT * WSF * U1) / H2;
        V2 = (QY1 - DTAV - m.DT * WSF * V1) / H2;
        const R cap = (m.variant == 1) ? (R)5.0 : (R)15.0;
        U2 = std::copysign(std::fmin(std::fabs(U2), cap), U2);
        V2 = std::copysign(std::fmin(std::fabs(V2), cap), V2);
    }
    W2 = sqrt_r<R>(U2 * U2 + V2 * V2);
    m.H[pos] = H2;
    m.U[pos] = U2;
    m.V[pos] = V2;
    m.Z[pos] = Z2;
    m.W[pos] = W2;
}

// Threads of the slot and cell loops (oracle_set_threads; 1 by default).
// Every slot writes only its own FLUX and KLAS entries and reads states the
// loop does not write, every cell only its own state, so any thread count
// gives the same bits.
int g_threads = 1;

template <class R> void calc_flux_all(const oracle_mesh &om, int kt, int jt) {
    View<R> m(om);
    const int n = 4 * m.CELL;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int idx = 0; idx < n; idx++) flux_slot<R>(m, kt, jt, idx);
}

template <class R> void update_all(const oracle_mesh &om) {
    View<R> m(om);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int pos = 0; pos < m.CELL; pos++) update_one<R>(m, pos);
}

int steps_per_day_of(const oracle_mesh &om) { return (int)(om.MDT / om.DT); }

}  // namespace

extern "C" {

int oracle_abi_version(void) { return 1; }

// Threads for the step loops (>= 1); returns the previous count.
int oracle_set_threads(int n) {
    const int old = g_threads;
    g_threads = n < 1 ? 1 : n;
    return old;
}

void oracle_calculate_flux(const oracle_mesh *m, int kt, int jt) {
    if (m->dtype == 1) calc_flux_all<float>(*m, kt, jt);
    else calc_flux_all<double>(*m, kt, jt);
}

void oracle_update_cell(const oracle_mesh *m) {
    if (m->dtype == 1) update_all<float>(*m);
    else update_all<double>(*m);
}

// Run nsteps steps in the order of cuda_native_impl/benchmark.cu:87-96,
// starting at (*day, *kt) and leaving the cursor after the last step.
// Returns the number of steps run (fewer than asked at the NDAYS cap).
int oracle_step_n(const oracle_mesh *m, int nsteps, int *day, int *kt) {
    const int spd = steps_per_day_of(*m);
    int s = 0;
    while (s < nsteps && *day < m->NDAYS) {
        if (*kt >= spd) { (*day)++; *kt = 1; continue; }
        oracle_calculate_flux(m, *kt, *day);
        oracle_update_cell(m);
        s++;
        (*kt)++;
    }
    if (*kt >= spd && *day < m->NDAYS) { (*day)++; *kt = 1; }
    return s;
}

// Test hook: one Osher solve (F:570-712) in double or float.
void oracle_osher(int dtype, const double QL[3], const double QR[3], double FIL, double H_pre,
                  double out[4]) {
    if (dtype == 1) {
        float ql[3] = {(float)QL[0], (float)QL[1], (float)QL[2]};
        float qr[3] = {(float)QR[0], (float)QR[1], (float)QR[2]};
        float o[4];
        osher<float>((float)H_pre, ql, qr, (float)FIL, o);
        for (int i = 0; i < 4; i++) out[i] = o[i];
    } else {
        osher<double>(H_pre, QL, QR, FIL, out);
    }
}

// Geometry preprocessing, restated from MeshData::preCalculate
// (cuda_native_impl/hydro-cal-src/src/mesh.cpp:280-372), fp64.
//   NAP   [4*CELL] 1-based node ids, slot-major as nap[4*i + j]
//   XP,YP [NOD]    node coordinates (shifted in place by their minima)
//   ZBC, Z, FNC0 [CELL]; Z is the initial level and is rewritten for dry cells
// Outputs: SIDE SINF COSF SLCOS SLSIN [4*CELL]; AREA ZB1 H FNC [CELL].
void oracle_precalculate(int CELL, int NOD, const int32_t *NAP, double *XP, double *YP,
                         const double *ZBC, double *Z, const double *FNC0, double HM1,
                         double *SIDE, double *SINF, double *COSF, double *SLCOS,
                         double *SLSIN, double *AREA, double *ZB1, double *H, double *FNC) {
    double xmin = XP[0], ymin = YP[0];
    for (int i = 1; i < NOD; i++) {
        if (XP[i] < xmin) xmin = XP[i];
        if (YP[i] < ymin) ymin = YP[i];
    }
    for (int i = 0; i < NOD; i++) {
        XP[i] -= xmin;
        YP[i] -= ymin;
    }
    for (int i = 0; i < 4 * CELL; i++) SIDE[i] = SINF[i] = COSF[i] = SLCOS[i] = SLSIN[i] = 0.0;
    for (int i = 0; i < CELL; i++) AREA[i] = 0.0;
    for (int i = 0; i < CELL; i++) {
        const int32_t *nw = NAP + 4 * i;
        if (nw[0] == 0) continue;
        int nv = (nw[3] == 0 || nw[3] == nw[0]) ? 3 : 4;
        double X1 = XP[nw[0] - 1], X2 = XP[nw[1] - 1], X3 = XP[nw[2] - 1];
        double Y1 = YP[nw[0] - 1], Y2 = YP[nw[1] - 1], Y3 = YP[nw[2] - 1];
        AREA[i] = std::fabs(((Y3 - Y1) * (X2 - X1) - (X3 - X1) * (Y2 - Y1)) / 2.0);
        if (nv == 4) {
            double X4 = XP[nw[3] - 1], Y4 = YP[nw[3] - 1];
            AREA[i] += std::fabs(((Y4 - Y1) * (X3 - X1) - (X4 - X1) * (Y3 - Y1)) / 2.0);
        }
        for (int j = 0; j < nv; j++) {
            int n1 = nw[j] - 1, n2 = nw[(j + 1) % nv] - 1;
            double DX = XP[n1] - XP[n2];
            double DY = YP[n2] - YP[n1];
            int s = 4 * i + j;
            SIDE[s] = std::sqrt(DX * DX + DY * DY);
            if (SIDE[s] > 0.0) {
                SINF[s] = DX / SIDE[s];
                COSF[s] = DY / SIDE[s];
            }
            SLCOS[s] = SIDE[s] * COSF[s];
            SLSIN[s] = SIDE[s] * SINF[s];
        }
    }
    for (int i = 0; i < CELL; i++) ZB1[i] = ZBC[i] + HM1;
    for (int i = 0; i < CELL; i++) {
        H[i] = 0.0;
        if (NAP[4 * i] == 0) continue;
        if (Z[i] <= ZBC[i]) {
            H[i] = HM1;
            Z[i] = ZB1[i];
        } else {
            H[i] = Z[i] - ZBC[i];
        }
    }
    for (int i = 0; i < CELL; i++) FNC[i] = 9.81 * FNC0[i] * FNC0[i];
}

}  // extern "C"
