// adjoint.hpp -- hand-written reverse mode of the fp64 timestep (variant A).
//
// For one slot flux F(own, nbr) (physics.hpp slot_flux) and one cell update,
// these functions recompute the forward values they need and then propagate
// an adjoint seed backwards, statement by statement, to the inputs.  Branches
// are taken exactly as in the forward step; min/max/abs/copysign
// differentiate the operand that was selected (the first one on ties), which
// is the same convention as the dual number of dual.hpp, so the adjoint and
// the tangent-linear model are transposes of each other.
//
// Supported slot kinds: interior (primary / mirror / wet-dry ladder), KLAS 4
// walls, KLAS 10 discharge, KLAS 1 water level.  The fixed-point iterations of
// KLAS 10 / KLAS 1 are differentiated through their iterations (forward mode
// with a few tangents; boundary slots are rare).
#pragma once
#include "physics.hpp"

namespace hydro {

// Inputs of one side of a slot: state and bed.
struct SideIn {
    double H, U, V, Z, ZB;
};
struct SideBar {
    double H = 0, U = 0, V = 0, Z = 0, ZB = 0;
};

// Adjoint of QF(h, u, v) accumulated with sign: acc += sign * (hu, hu u, hu v, 4.905 h h)
__host__ __device__ __forceinline__ void qf_bar(const double g[4], bool neg, double h, double u, double v,
                                                double &hb, double &ub, double &vb) {
    const double s = neg ? -1.0 : 1.0;
    const double g0 = s * g[0], g1 = s * g[1], g2 = s * g[2], g3 = s * g[3];
    const double f0 = h * u;
    const double gf0 = g0 + g1 * u + g2 * v;
    hb += gf0 * u + g3 * 9.81 * h;
    ub += gf0 * h + g1 * f0;
    vb += g2 * f0;
}

// Reverse of osher() (physics.hpp).  g = adjoint of out[4].  Adds to the
// adjoints of (H_pre == QL0 here), QL1, QL2, QR0, QR1, QR2, FIL.
__host__ __device__ inline void osher_bar(double H_pre, double QL1, double QL2, double QR0, double QR1,
                                          double QR2, double FIL, const double g[4], double &Hb, double &QL1b,
                                          double &QL2b, double &QR0b, double &QR1b, double &QR2b,
                                          double &FILb) {
    const double CR = sqrt(9.81 * QR0);
    const double FIR = QR1 - 2 * CR;
    const double UA = (FIL + FIR) / 2;
    const double CA = fabs((FIL - FIR) / 4);
    const double CL = sqrt(9.81 * H_pre);
    const int K2 = (CA < UA) ? 1 : (UA >= 0.0 && UA < CA) ? 2 : (UA >= -CA && UA < 0.0) ? 3 : 4;
    const bool rsub = (QR1 >= -CR);
    const int K1 = (QL1 < CL && rsub) ? 1 : (QL1 >= CL && rsub) ? 2 : (QL1 < CL && QR1 < -CR) ? 3 : 4;
    // The same single bit tests as the forward osher(): the gfx950 backend
    // of this toolchain mis-lowers the two-case switch that
    // (K1 == 2 || K1 == 4) becomes.
    const bool lsup = (K1 & 1) == 0;
    const bool rneg = K1 >= 3;
    double FIRb = 0, UAb = 0, FIRmb = 0;
    const bool useM = (K2 == 2 || K2 == 3);
    const bool mleft = (K2 == 2);
    const double X = useM ? ((mleft ? FIL : FIR) - UA) : 0.0;
    const double FIRm = (useM && !mleft) ? X : FIR;
    // right state
    if (rneg) qf_bar(g, false, QR0, QR1, QR2, QR0b, QR1b, QR2b);
    // right sonic point: US = FIRm / 3, HS = US^2 / 9.81
    if (rneg ? (K2 != 4) : (K2 == 4)) {
        const double US = FIRm / 3;
        const double HS = US * US / 9.81;
        double hb = 0, ub = 0;
        qf_bar(g, rneg, HS, US, QR2, hb, ub, QR2b);
        FIRmb += (ub + hb * 2.0 * US / 9.81) / 3.0;
    }
    // intermediate state
    if (useM) {
        const double HA = X * X / 39.24;
        double hb = 0, ub = 0, vb = 0;
        qf_bar(g, false, HA, UA, mleft ? QL2 : QR2, hb, ub, vb);
        if (mleft) QL2b += vb;
        else QR2b += vb;
        UAb += ub;
        double Xb = hb * 2.0 * X / 39.24;
        if (!mleft) Xb += FIRmb;     // FIRm is X after QS<5>
        else FIRb += FIRmb;
        if (mleft) FILb += Xb;
        else FIRb += Xb;
        UAb -= Xb;
    } else {
        FIRb += FIRmb;
    }
    // left sonic point: US = FIL / 3
    if (lsup ? (K2 != 1) : (K2 == 1)) {
        const double US = FIL / 3;
        const double HS = US * US / 9.81;
        double hb = 0, ub = 0;
        qf_bar(g, lsup, HS, US, QL2, hb, ub, QL2b);
        FILb += (ub + hb * 2.0 * US / 9.81) / 3.0;
    }
    // left state (QL0 == H_pre)
    if (lsup) qf_bar(g, false, H_pre, QL1, QL2, Hb, QL1b, QL2b);
    // UA = (FIL + FIR) / 2 ; FIR = QR1 - 2 CR ; CR = sqrt(9.81 QR0)
    FILb += UAb / 2;
    FIRb += UAb / 2;
    QR1b += FIRb;
    const double CRb = -2.0 * FIRb;
    if (CR > 0) QR0b += CRb * 9.81 / (2.0 * CR);
}

// ---------------------------------------------------------------------------
// Small forward-mode helper for the boundary fixed-point iterations.
template <int N> struct Tan {
    double v, d[N];
};

// KLAS 10 (F:204-232): HB as a function of (FIL, H1); returns HB and dHB/d(FIL, H1).
__host__ __device__ inline double klas10_hb(double F0, double FIL, double H1, double dHB[2]) {
    double HB0 = H1, dHB0[2] = {0, 1};
    double HB = 0, dh[2] = {0, 0};
    for (int k = 1; k <= 20; k++) {
        const double Wt = FIL - F0 / HB0;
        double dW[2];
        for (int i = 0; i < 2; i++) dW[i] = (i == 0 ? 1.0 : 0.0) + F0 / (HB0 * HB0) * dHB0[i];
        HB = Wt * Wt / 39.24;
        for (int i = 0; i < 2; i++) dh[i] = 2.0 * Wt * dW[i] / 39.24;
        if (fabs(HB0 - HB) <= 0.005) break;
        HB0 = HB0 * 0.5 + HB * 0.5;
        for (int i = 0; i < 2; i++) dHB0[i] = 0.5 * dHB0[i] + 0.5 * dh[i];
    }
    dHB[0] = dh[0];
    dHB[1] = dh[1];
    return HB;
}

// URB iteration (F:379-388) as a function of (QLn, H_pre, HB1).
__host__ __device__ inline double urb_tan(double QLn, double H_pre, double HB1, double tol, double dU[3]) {
    const double sH = sqrt(H_pre), sB = sqrt(HB1);
    const double FIAL = QLn + 6.264 * sH;
    const double dFIAL[3] = {1.0, sH > 0 ? 6.264 / (2 * sH) : 0.0, 0.0};
    double UR0 = QLn, dUR0[3] = {1, 0, 0};
    double URB = UR0, dURB[3] = {1, 0, 0};
    for (int it = 1; it <= 30; it++) {
        const double FIAR = URB - 6.264 * sB;
        double dFIAR[3];
        for (int i = 0; i < 3; i++) dFIAR[i] = dURB[i] - (i == 2 && sB > 0 ? 6.264 / (2 * sB) : 0.0);
        const double a = FIAL + FIAR, b = FIAL - FIAR;
        const double num = a * b * b;
        URB = num / HB1 / 313.92;
        for (int i = 0; i < 3; i++) {
            const double da = dFIAL[i] + dFIAR[i], db = dFIAL[i] - dFIAR[i];
            const double dnum = da * b * b + 2.0 * a * b * db;
            dURB[i] = (dnum / HB1 - (i == 2 ? num / (HB1 * HB1) : 0.0)) / 313.92;
        }
        if (fabs(URB - UR0) <= tol) break;
        UR0 = URB;
        for (int i = 0; i < 3; i++) dUR0[i] = dURB[i];
    }
    for (int i = 0; i < 3; i++) dU[i] = dURB[i];
    return URB;
}

// ---------------------------------------------------------------------------
// Reverse of slot_flux<double, 0> for one slot.  Fb = adjoint of (F0..F3).
// Returns false for slot kinds without an adjoint.  BcT is BcCtx<double>, or
// for an fp32 trajectory a context whose boundary series are the fp32 arrays
// (their values enter the recomputation as they are stored).
template <class BcT>
__host__ __device__ inline bool slot_flux_bar(int kind, const Scalars<double> &S, const BcT &bc,
                                              const SideIn &me, bool has, const SideIn &nb, double COSJ,
                                              double SINJ, const double Fb[4], SideBar &meb, SideBar &nbb) {
    using K = Consts<double>;
    const double H1 = me.H, U1 = me.U, V1 = me.V, Z1 = me.Z, ZBp = me.ZB;
    const double ZB1 = ZBp + S.HM1;
    const double QL1 = U1 * COSJ + V1 * SINJ;
    const double QL2 = V1 * COSJ - U1 * SINJ;
    const double ZI = fmax(Z1, ZB1);
    double HC = 0, BC = 0, ZC = 0, UC = 0, VC = 0;
    if (has) {
        HC = fmax(nb.H, S.HM1);
        BC = nb.ZB;
        ZC = fmax(BC, nb.Z);
        UC = nb.U;
        VC = nb.V;
    }
    double H1b = 0, QL1b = 0, QL2b = 0, ZIb = 0, HCb = 0, BCb = 0, ZCb = 0, UCb = 0, VCb = 0, ZBpb = 0;
    if (kind >= K_B10) {
        const double CL = sqrt(9.81 * H1);
        const double FIL = QL1 + 2 * CL;
        if (QL1 > CL && H1 < S.HM2) {
            const double F0 = H1 * QL1;
            const double gF0 = Fb[0] + Fb[1] * QL1 + Fb[2] * QL2;
            H1b += gF0 * QL1 + Fb[3] * 9.81 * H1;
            QL1b += gF0 * H1 + Fb[1] * F0;
            QL2b += Fb[2] * F0;
        } else {
            // F2 = QL1 > 0 ? H1*QL1*QL2 : 0 unless the kind overwrites it
            // kinds {B1, B3, B5, B7, BUNK} as one mask test (no multi-case switch)
            constexpr unsigned kF2Pre = (1u << K_B1) | (1u << K_B3) | (1u << K_B5) | (1u << K_B7) | (1u << K_BUNK);
            const bool f2pre = (kF2Pre >> kind) & 1u;
            if (f2pre && QL1 > 0) {
                H1b += Fb[2] * QL1 * QL2;
                QL1b += Fb[2] * H1 * QL2;
                QL2b += Fb[2] * H1 * QL1;
            }
            if (kind == K_B4) {
                H1b += Fb[3] * 9.81 * H1;
            } else if (kind == K_B10) {
                const size_t row = (size_t)bc.jt * (size_t)bc.CELL + (size_t)bc.cell;
                double F0 = -(bc.QT[row] + bc.DQT[row] * bc.kt);
                F0 = F0 / bc.SIDE;
                const double QB2 = F0 * F0;
                double dHB[2];
                const double HB = klas10_hb(F0, FIL, H1, dHB);
                double HBb = Fb[3] * 9.81 * HB;
                if (!(HB <= 1)) HBb += -Fb[1] * QB2 / (HB * HB);
                QL1b += HBb * dHB[0];                     // dFIL/dQL1 = 1
                H1b += HBb * dHB[0] * (CL > 0 ? 9.81 / CL : 0.0) + HBb * dHB[1];
            } else if (kind == K_B1) {
                const size_t row = (size_t)bc.jt * (size_t)bc.CELL + (size_t)bc.cell;
                const double HB1r = bc.ZT[row] + bc.DZT[row] * bc.jt - ZBp;
                const double HB1 = fmax(HB1r, S.HM2);
                double dU[3];
                const double URB = urb_tan(QL1, H1, HB1, 0.0001, dU);
                const double F0 = HB1 * URB;
                // F0 = HB1 URB ; F1 = F0 URB ; F3 = 4.905 HB1^2
                const double gF0 = Fb[0] + Fb[1] * URB;
                const double URBb = gF0 * HB1 + Fb[1] * F0;
                double HB1b = gF0 * URB + Fb[3] * 9.81 * HB1 + URBb * dU[2];
                QL1b += URBb * dU[0];
                H1b += URBb * dU[1];
                if (HB1r >= S.HM2) ZBpb -= HB1b;
            } else {
                return false;
            }
        }
    } else if (H1 <= S.HM1 && HC <= S.HM1) {
        // zero flux
    } else if (ZI <= BC) {
        // F0 = -C1 HC^1.5 ; F1 = H1 QL1 |QL1| ; F3 = 4.905 H1^2
        HCb += Fb[0] * (-(double)K::C1) * 1.5 * sqrt(HC);
        H1b += Fb[1] * QL1 * fabs(QL1) + Fb[3] * 9.81 * H1;
        QL1b += Fb[1] * H1 * 2.0 * fabs(QL1);
    } else if (ZC <= ZBp) {
        // F0 = C1 H1^1.5 ; F1 = H1 |QL1| QL1 ; F2 = H1 |QL1| QL2
        const double aq = fabs(QL1), sg = QL1 < 0 ? -1.0 : 1.0;
        H1b += Fb[0] * (double)K::C1 * 1.5 * sqrt(H1) + Fb[1] * aq * QL1 + Fb[2] * aq * QL2;
        QL1b += Fb[1] * H1 * 2.0 * aq + Fb[2] * H1 * sg * QL2;
        QL2b += Fb[2] * H1 * aq;
    } else if (H1 <= S.HM2) {
        if (ZC > ZI) {
            const double DHr = ZC - ZBp;
            const double DH = fmax(DHr, S.HM1);
            const double sd = sqrt(DH);
            const double UN = -(double)K::C1 * sd;
            const double F0 = DH * UN;
            const double t = VC * COSJ - UC * SINJ;
            const double gF0 = Fb[0] + Fb[1] * UN + Fb[2] * t;
            const double UNb = gF0 * DH + Fb[1] * F0;
            const double DHb = gF0 * UN + UNb * (-(double)K::C1) / (2.0 * sd);
            const double tb = Fb[2] * F0;
            VCb += tb * COSJ;
            UCb -= tb * SINJ;
            if (DHr >= S.HM1) {
                ZCb += DHb;
                ZBpb -= DHb;
            }
        } else {
            H1b += Fb[0] * (double)K::C1 * 1.5 * sqrt(H1);
        }
        H1b += Fb[3] * 9.81 * H1;
    } else if (HC <= S.HM2) {
        if (ZI > ZC) {
            const double DHr = ZI - BC;
            const double DH = fmax(DHr, S.HM1);
            const double sd = sqrt(DH);
            const double UN = (double)K::C1 * sd;
            const double HC1 = ZC - ZBp;
            const double F0 = DH * UN;
            const double gF0 = Fb[0] + Fb[1] * UN + Fb[2] * QL2;
            const double UNb = gF0 * DH + Fb[1] * F0;
            const double DHb = gF0 * UN + UNb * (double)K::C1 / (2.0 * sd);
            QL2b += Fb[2] * F0;
            const double HC1b = Fb[3] * 9.81 * HC1;
            ZCb += HC1b;
            ZBpb -= HC1b;
            if (DHr >= S.HM1) {
                ZIb += DHb;
                BCb -= DHb;
            }
        } else {
            HCb += Fb[0] * (-(double)K::C1) * 1.5 * sqrt(HC);
            H1b += Fb[1] * QL1 * QL1 + Fb[3] * 9.81 * H1;
            QL1b += Fb[1] * H1 * 2.0 * QL1;
        }
    } else {
        // both sides wet: Riemann problem on (left a, right b)
        const bool prim = (kind == K_PRIMARY);
        const double c = prim ? COSJ : -COSJ, s = prim ? SINJ : -SINJ;
        const double Ha = prim ? H1 : nb.H, Ua = prim ? U1 : nb.U, Va = prim ? V1 : nb.V;
        const double ZBa = prim ? ZBp : nb.ZB;
        const double Hb = prim ? nb.H : H1, Ub = prim ? nb.U : U1, Vb = prim ? nb.V : V1;
        const double ZBb = prim ? nb.ZB : ZBp, Zb = prim ? nb.Z : Z1;
        const double qa1 = Ua * c + Va * s, qa2 = Va * c - Ua * s;
        const double CLa = sqrt(9.81 * Ha);
        const double FILa = qa1 + 2 * CLa;
        const double ZCbv = fmax(ZBb, Zb);
        const double HCbv = fmax(Hb, S.HM1);
        const double QR0r = ZCbv - ZBa;
        const double QR0 = fmax(QR0r, S.HM1);
        const double UR = Ub * c + Vb * s;
        const double r0 = HCbv / QR0;
        const double ratio = fmin(r0, 1.5);
        const bool shallow = (HCbv <= S.HM2 || QR0 <= S.HM2);
        const double QR1 = shallow ? copysign((double)K::VMIN, UR) : UR * ratio;
        const double QR2 = Vb * c - Ub * s;
        double FO[4];
        osher<double>(Ha, Ha, qa1, qa2, QR0, QR1, QR2, FILa, FO);
        double g[4];
        double ratiob = 0, HCbb = 0, URb = 0, ZBab = 0, ZBpb_m = 0;
        // F1 = FO1 + (1 - ratio) HCb UR UR / 2 ; F2 = FO2
        g[1] = Fb[1];
        g[2] = Fb[2];
        ratiob += -Fb[1] * HCbv * UR * UR / 2;
        HCbb += Fb[1] * (1 - ratio) * UR * UR / 2;
        URb += Fb[1] * (1 - ratio) * HCbv * UR;
        if (prim) {
            g[0] = Fb[0];
            g[3] = Fb[3];
        } else {
            g[0] = -Fb[0];
            const double q = FO[3] / 4.905;
            const double sq = sqrt(q);
            const double ZA = sq + ZBa;
            g[3] = 0;
            if (ZA > ZBp) {
                const double HC3 = ZA - ZBp;
                const double ZAb = Fb[3] * 9.81 * HC3;
                ZBpb_m -= ZAb;
                ZBab += ZAb;
                if (sq > 0) g[3] = ZAb / (2.0 * sq) / 4.905;
            }
        }
        double Hab = 0, qa1b = 0, qa2b = 0, QR0b = 0, QR1b = 0, QR2b = 0, FILb = 0;
        osher_bar(Ha, qa1, qa2, QR0, QR1, QR2, FILa, g, Hab, qa1b, qa2b, QR0b, QR1b, QR2b, FILb);
        double Ubb = 0, Vbb = 0, Hbb = 0, Zbb = 0, ZBbb = 0;
        Vbb += QR2b * c;
        Ubb -= QR2b * s;
        if (!shallow) {
            URb += QR1b * ratio;
            ratiob += QR1b * UR;
        }
        if (r0 <= 1.5) {
            HCbb += ratiob / QR0;
            QR0b -= ratiob * HCbv / (QR0 * QR0);
        }
        if (QR0r >= S.HM1) {
            // QR0 = ZCb - ZBa ; ZCb = fmax(ZBb, Zb)
            if (ZBb >= Zb) ZBbb += QR0b;
            else Zbb += QR0b;
            ZBab -= QR0b;
        }
        if (Hb >= S.HM1) Hbb += HCbb;
        Ubb += URb * c;
        Vbb += URb * s;
        // FIL = qa1 + 2 CLa ; CLa = sqrt(9.81 Ha)
        qa1b += FILb;
        if (CLa > 0) Hab += 2.0 * FILb * 9.81 / (2.0 * CLa);
        double Uab = qa1b * c - qa2b * s, Vab = qa1b * s + qa2b * c;
        // map (a, b) back to (me, nbr)
        SideBar &A = prim ? meb : nbb, &B = prim ? nbb : meb;
        A.H += Hab;
        A.U += Uab;
        A.V += Vab;
        A.ZB += ZBab;
        B.H += Hbb;
        B.U += Ubb;
        B.V += Vbb;
        B.Z += Zbb;
        B.ZB += ZBbb;
        meb.ZB += ZBpb_m;
        return true;
    }
    // map the ladder / boundary intermediates back to the inputs
    meb.H += H1b;
    meb.U += QL1b * COSJ - QL2b * SINJ;
    meb.V += QL1b * SINJ + QL2b * COSJ;
    meb.ZB += ZBpb;
    // ZI = fmax(Z1, ZB1), ZB1 = ZBp + HM1
    if (Z1 >= ZB1) meb.Z += ZIb;
    else meb.ZB += ZIb;
    if (has) {
        if (nb.H >= S.HM1) nbb.H += HCb;
        nbb.ZB += BCb;
        if (BC >= nb.Z) nbb.ZB += ZCb;
        else nbb.Z += ZCb;
        nbb.U += UCb;
        nbb.V += VCb;
    }
    return true;
}

// ---------------------------------------------------------------------------
// Reverse of the cell update (variant A, update_math<double, 0>).
struct UpdateBar {
    double WH = 0, WU = 0, WV = 0;  // adjoints of the flux sums
    double H1 = 0, U1 = 0, V1 = 0;  // direct adjoints of the old state
    double FNC = 0, ZBC = 0;
};

__host__ __device__ inline UpdateBar update_bar(const Scalars<double> &S, double H1, double U1, double V1,
                                                double AREA, double FNC, double WH, double WU, double WV,
                                                double lH, double lU, double lV, double lZ) {
    using K = Consts<double>;
    UpdateBar b;
    const double DTA = 1.0 * S.DT / (1.0 * AREA);
    const double WDTA = 1.00 * DTA;
    const double H2r = H1 - WDTA * WH + K::QLUA;
    const double H2 = fmax(H2r, S.HM1);
    double H2b = lH + lZ;
    b.ZBC += lZ;
    if (H2 <= S.HM1) {
        // U2 = V2 = 0
    } else if (H2 <= S.HM2) {
        if (fabs(U1) < K::VMIN) b.U1 += lU;
        if (fabs(V1) < K::VMIN) b.V1 += lV;
    } else {
        const double sp = sqrt(U1 * U1 + V1 * V1);
        const double ph = pow_d(H1, 0.33333);
        const double WSF = FNC * sp / ph;
        const double NU = H1 * U1 - WDTA * WU - S.DT * WSF * U1;
        const double NV = H1 * V1 - WDTA * WV - S.DT * WSF * V1;
        const double U2r = NU / H2, V2r = NV / H2;
        const double U2b = fabs(U2r) < 15.0 ? lU : 0.0;
        const double V2b = fabs(V2r) < 15.0 ? lV : 0.0;
        const double NUb = U2b / H2, NVb = V2b / H2;
        H2b -= U2b * U2r / H2 + V2b * V2r / H2;
        b.H1 += NUb * U1 + NVb * V1;
        b.U1 += NUb * (H1 - S.DT * WSF);
        b.V1 += NVb * (H1 - S.DT * WSF);
        b.WU -= NUb * WDTA;
        b.WV -= NVb * WDTA;
        const double WSFb = -(NUb * S.DT * U1 + NVb * S.DT * V1);
        b.FNC += WSFb * sp / ph;
        const double spb = WSFb * FNC / ph;
        const double phb = -WSFb * WSF / ph;
        if (sp > 0) {
            b.U1 += spb * U1 / sp;
            b.V1 += spb * V1 / sp;
        }
        b.H1 += phb * 0.33333 * ph / H1;
    }
    if (H2r >= S.HM1) {
        b.H1 += H2b;
        b.WH -= H2b * WDTA;
    }
    return b;
}

// Adjoint of the flux sums -> adjoint of one slot's flux (F:732-742).
__host__ __device__ __forceinline__ void flux_bar_of_sums(const UpdateBar &u, double SL, double SLCA, double SLSA,
                                                          double Fb[4]) {
    Fb[0] = SL * u.WH;
    Fb[1] = SLCA * u.WU + SLSA * u.WV;
    Fb[2] = -SLSA * u.WU + SLCA * u.WV;
    Fb[3] = Fb[1];
}

}  // namespace hydro
