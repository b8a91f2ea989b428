// Device-vs-host check of the timestep's device functions
// (hydro-bench_amd/csrc/physics.hpp, pow_fixed.hpp, and the reverse-mode
// functions of adjoint.hpp): random inputs covering every Osher case, every
// wet/dry ladder branch and every boundary kind are
// evaluated by a gfx950 kernel and by the same source compiled for the host,
// and compared.  It guards against the device compiler lowering a construct
// differently from its host compile (the host build is itself compared with
// the oracle by the CPU tests).  Test infrastructure only, built into
// tests/native/_build/libdevice_check.so; the Python test loads it after
// torch so both share one HIP runtime.
//
//   int dc_run(int group, int n, uint64_t seed, double *max_ulp, int64_t *inexact, int64_t *bad)
// runs n cases of `group` and reports the largest difference in ulps of the
// result type, the number of results that are not bit-identical, and the
// number differing by more than 4 ulps (for the cell update: and by more
// than an absolute floor, see dc_run); NaN matches NaN.  Returns 0, or a
// HIP error code.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "adjoint.hpp"
#include "physics.hpp"

using namespace hydro;

namespace {

constexpr int NX = 24;   // inputs per case
constexpr int NO = 10;   // outputs per case
constexpr int NDAYS = 2, NHQ = 4;

// Boundary series of the one-cell boundary context every boundary case uses.
struct BcArrays {
    const double *QT, *DQT, *ZT, *DZT, *ZW, *QW, *BF;
    const float *QTf, *DQTf, *ZTf, *DZTf, *ZWf, *QWf, *BFf;
    const int32_t *NHQ1;
};

enum Group {
    G_OSHER = 0, G_SLOT64_A, G_SLOT64_B, G_PAIR64, G_UPDATE64, G_SLOT32_A, G_SLOT32_B, G_PAIR32, G_UPDATE32, G_POW,
    G_OSHER_BAR, G_SLOT_BAR, G_UPDATE_BAR,
    G_COUNT
};

template <class R, class T>
__host__ __device__ BcCtx<R> make_ctx(const T *QT, const T *DQT, const T *ZT, const T *DZT, const T *ZW, const T *QW,
                                      const T *BF, const int32_t *NHQ1, const double *x) {
    BcCtx<R> bc;
    bc.QT = QT; bc.DQT = DQT; bc.ZT = ZT; bc.DZT = DZT; bc.ZW = ZW; bc.QW = QW; bc.BF = BF; bc.NHQ1 = NHQ1;
    bc.CELL = 1; bc.NHQ = NHQ; bc.cell = 0;
    bc.kt = (int)x[16];
    bc.jt = (int)x[17];
    bc.SIDE = (R)x[15];
    bc.COSF = (R)x[12];
    bc.SINF = (R)x[13];
    return bc;
}

// x: 0-5 own H U V Z ZB1 ZBC | 6-10 nbr H U V Z ZBC | 11 has | 12,13 c s |
//    14 prev (all four) | 15 SIDE | 16 kt | 17 jt | 18 AREA | 19 FNC |
//    20-23 adjoint seeds of the four outputs (reverse-mode groups)
template <class R, int VAR>
__host__ __device__ void eval_slot(const double *x, int kind, const BcCtx<R> &bc, double *o) {
    const Scalars<R> S{(R)1.0, (R)0.001, (R)0.01, (R)0.0};
    bool fired = false;
    const R pv = (R)x[14];
    const Flux<R> f = slot_flux<R, VAR, true>(kind, S, bc, (R)x[0], (R)x[1], (R)x[2], (R)x[3], (R)x[4], (R)x[5],
                                              x[11] != 0, (R)x[6], (R)x[7], (R)x[8], (R)x[9], (R)x[10], (R)x[12],
                                              (R)x[13], Flux<R>{pv, pv, pv, pv}, &fired);
    o[0] = f.f0; o[1] = f.f1; o[2] = f.f2; o[3] = f.f3; o[4] = fired ? 1 : 0;
}

template <class R> __host__ __device__ void eval_pair(const double *x, double *o) {
    const Scalars<R> S{(R)1.0, (R)0.001, (R)0.01, (R)0.0};
    Flux<R> a, b;
    pair_flux<R>(S, (R)x[0], (R)x[1], (R)x[2], (R)x[3], (R)x[4], (R)x[5], (R)x[6], (R)x[7], (R)x[8], (R)x[9],
                 (R)x[10] + (R)0.001, (R)x[10], (R)x[12], (R)x[13], &a, &b);
    o[0] = a.f0 + 3 * b.f0;   // both sides in one comparison
    o[1] = a.f1; o[2] = b.f1; o[3] = a.f3 - b.f3; o[4] = a.f2 + b.f2;
}

template <class R, int VAR> __host__ __device__ void eval_update(const double *x, double *o) {
    const Scalars<R> S{(R)1.0, (R)0.001, (R)0.01, (R)0.0};
    // sides: positive lengths, a fourth one for quads only
    const CellOut<R> c = update_math<R, VAR>(S, (R)fmax(x[0], 0.001), (R)x[1], (R)x[2], (R)x[5], (R)x[18],
                                             (R)x[19], (R)x[6], (R)x[7], (R)x[8], (R)x[15],
                                             (R)(5 + 30 * fabs(x[12])), (R)(5 + 30 * fabs(x[13])),
                                             x[11] != 0 ? (R)(5 + 10 * fabs(x[14])) : (R)0);
    o[0] = c.H; o[1] = c.U; o[2] = c.V; o[3] = c.Z; o[4] = c.W;
}

// Reverse of the Osher flux: the seven input adjoints for seeds x[20..23].
__host__ __device__ void eval_osher_bar(const double *x, double *o) {
    const double g[4] = {x[20], x[21], x[22], x[23]};
    double Hb = 0, QL1b = 0, QL2b = 0, QR0b = 0, QR1b = 0, QR2b = 0, FILb = 0;
    osher_bar(x[0], x[1], x[2], x[6], x[7], x[8], x[1] + 2 * sqrt(9.81 * x[0]), g, Hb, QL1b, QL2b, QR0b, QR1b,
              QR2b, FILb);
    o[0] = Hb; o[1] = QL1b; o[2] = QL2b; o[3] = QR0b; o[4] = QR1b; o[5] = QR2b; o[6] = FILb;
}

// Reverse of one slot flux (kinds with an adjoint: interior, KLAS 10, 1, 4):
// the adjoints of both sides' H U V Z ZB.
__host__ __device__ void eval_slot_bar(const double *x, int kind, const BcCtx<double> &bc, double *o) {
    const Scalars<double> S{1.0, 0.001, 0.01, 0.0};
    const SideIn me{x[0], x[1], x[2], x[3], x[5]}, nb{x[6], x[7], x[8], x[9], x[10]};
    const double Fb[4] = {x[20], x[21], x[22], x[23]};
    SideBar meb, nbb;
    const bool ok = slot_flux_bar(kind, S, bc, me, x[11] != 0, nb, x[12], x[13], Fb, meb, nbb);
    o[0] = meb.H; o[1] = meb.U; o[2] = meb.V; o[3] = meb.Z; o[4] = meb.ZB;
    o[5] = nbb.H; o[6] = nbb.U; o[7] = nbb.V; o[8] = nbb.Z; o[9] = ok ? nbb.ZB : -1.0;
}

// Reverse of the cell update (variant A): adjoints of the flux sums, the old
// state, FNC and ZBC for seeds x[20..23] on (H, U, V, Z).
__host__ __device__ void eval_update_bar(const double *x, double *o) {
    const Scalars<double> S{1.0, 0.001, 0.01, 0.0};
    const UpdateBar b = update_bar(S, fmax(x[0], 0.001), x[1], x[2], x[18], x[19], x[6], x[7], x[8], x[20], x[21],
                                   x[22], x[23]);
    o[0] = b.WH; o[1] = b.WU; o[2] = b.WV; o[3] = b.H1; o[4] = b.U1; o[5] = b.V1; o[6] = b.FNC; o[7] = b.ZBC;
}

__host__ __device__ void eval(int g, const double *x, int kind, const BcArrays &A, double *o) {
    for (int k = 0; k < NO; k++) o[k] = 0;
    switch (g) {
    case G_OSHER: {
        double FO[4];
        osher<double>(x[0], x[0], x[1], x[2], x[6], x[7], x[8], x[1] + 2 * sqrt(9.81 * x[0]), FO);
        o[0] = FO[0]; o[1] = FO[1]; o[2] = FO[2]; o[3] = FO[3]; o[4] = 0;
    } break;
    case G_SLOT64_A:
        eval_slot<double, 0>(x, kind, make_ctx<double>(A.QT, A.DQT, A.ZT, A.DZT, A.ZW, A.QW, A.BF, A.NHQ1, x), o);
        break;
    case G_SLOT64_B:
        eval_slot<double, 1>(x, kind, make_ctx<double>(A.QT, A.DQT, A.ZT, A.DZT, A.ZW, A.QW, A.BF, A.NHQ1, x), o);
        break;
    case G_PAIR64: eval_pair<double>(x, o); break;
    case G_UPDATE64:
        if (kind & 1) eval_update<double, 1>(x, o);
        else eval_update<double, 0>(x, o);
        break;
    case G_SLOT32_A:
        eval_slot<float, 0>(x, kind, make_ctx<float>(A.QTf, A.DQTf, A.ZTf, A.DZTf, A.ZWf, A.QWf, A.BFf, A.NHQ1, x), o);
        break;
    case G_SLOT32_B:
        eval_slot<float, 1>(x, kind, make_ctx<float>(A.QTf, A.DQTf, A.ZTf, A.DZTf, A.ZWf, A.QWf, A.BFf, A.NHQ1, x), o);
        break;
    case G_PAIR32: eval_pair<float>(x, o); break;
    case G_UPDATE32:
        if (kind & 1) eval_update<float, 1>(x, o);
        else eval_update<float, 0>(x, o);
        break;
    case G_OSHER_BAR: eval_osher_bar(x, o); break;
    case G_SLOT_BAR:
        eval_slot_bar(x, kind, make_ctx<double>(A.QT, A.DQT, A.ZT, A.DZT, A.ZW, A.QW, A.BF, A.NHQ1, x), o);
        break;
    case G_UPDATE_BAR: eval_update_bar(x, o); break;
    default:
        o[0] = pow1p5(x[0]); o[1] = pow_0p33333(x[0]); o[2] = pow1p5(x[1]); o[3] = pow_0p33333(x[1]);
        o[4] = pow_d(x[0], 2.0);
        break;
    }
}

__global__ void k_eval(int g, const double *x, const int *kind, BcArrays A, double *o, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) eval(g, x + (size_t)i * NX, kind[i], A, o + (size_t)i * NO);
}

// Inputs: mixtures of deep, shallow (near HM1 / HM2), dry and below-bed
// states; nearly level beds for the outflow branches; velocities up to
// supercritical; every slot kind.
void gen(int g, int n, uint64_t seed, std::vector<double> &X, std::vector<int> &K) {
    std::mt19937_64 r(seed);
    std::uniform_real_distribution<double> U(0, 1);
    X.assign((size_t)n * NX, 0);
    K.assign(n, 0);
    const int kinds[] = {K_PRIMARY, K_MIRROR, K_B10, K_B1, K_B3, K_B4, K_B5, K_B6, K_B7, K_BUNK};
    const int adj_kinds[] = {K_PRIMARY, K_MIRROR, K_B10, K_B1, K_B4};   // the kinds with an adjoint
    auto depth = [&]() {
        const double q = U(r);
        if (q < 0.12) return 0.001;
        if (q < 0.25) return 0.001 + 0.02 * U(r);
        if (q < 0.3) return 0.0;
        return 0.05 + 4.0 * U(r);
    };
    for (int i = 0; i < n; i++) {
        double *x = &X[(size_t)i * NX];
        if (g == G_POW) {
            x[0] = std::ldexp(0.5 + 0.5 * U(r), -60 + (int)(U(r) * 120));
            x[1] = std::pow(10.0, -6 + 10 * U(r));
            continue;
        }
        for (int k = 20; k < 24; k++) x[k] = 2 * (U(r) - 0.5);
        if (g == G_OSHER || g == G_OSHER_BAR) {
            x[0] = 0.01 + 4 * U(r); x[1] = 20 * (U(r) - 0.5); x[2] = 10 * (U(r) - 0.5);
            x[6] = 0.01 + 4 * U(r); x[7] = 20 * (U(r) - 0.5); x[8] = 10 * (U(r) - 0.5);
            continue;
        }
        const bool flat = (i % 5) == 1;
        const double zbP = 10.0 + 2.0 * (U(r) - 0.5);
        const double zbN = flat ? zbP + 0.01 * (U(r) - 0.5) : 10.0 + 2.0 * (U(r) - 0.5);
        const double hP = flat ? 0.02 * U(r) : depth(), hN = flat ? 0.02 * U(r) : depth();
        x[0] = hP;
        x[1] = 6 * (U(r) - 0.5) * ((i % 7 == 0) ? 4 : 1);
        x[2] = 6 * (U(r) - 0.5);
        x[3] = (i % 11 == 0) ? zbP - 0.1 * U(r) : zbP + hP;
        x[4] = zbP + 0.001;
        x[5] = zbP;
        x[6] = hN;
        x[7] = 6 * (U(r) - 0.5) * ((i % 7 == 0) ? 4 : 1);
        x[8] = 6 * (U(r) - 0.5);
        x[9] = (i % 13 == 0) ? zbN - 0.1 * U(r) : zbN + hN;
        x[10] = zbN;
        const double th = 2 * M_PI * U(r);
        x[12] = std::cos(th);
        x[13] = std::sin(th);
        x[14] = 2 * (U(r) - 0.5);
        x[15] = 5 + 45 * U(r);
        x[16] = 1 + (int)(U(r) * 59);
        x[17] = (int)(U(r) * NDAYS);
        x[18] = 50 + 500 * U(r);
        x[19] = 0.0005 + 0.01 * U(r);
        if (g == G_UPDATE64 || g == G_UPDATE32 || g == G_UPDATE_BAR) {
            x[6] = 50 * (U(r) - 0.5);      // WH
            x[7] = 200 * (U(r) - 0.5);     // WU
            x[8] = 200 * (U(r) - 0.5);     // WV
            x[11] = (i % 3) ? 1 : 0;       // quad (fourth side) or triangle
            K[i] = i & 1;                  // variant
            continue;
        }
        const int kind = (g == G_SLOT_BAR) ? adj_kinds[i % 5] : kinds[i % 10];
        K[i] = kind;
        x[11] = (kind <= K_MIRROR) ? ((i % 17) ? 1 : 0) : 0;
    }
}

int64_t ulp_diff(double a, double b, bool single) {
    if (std::isnan(a) && std::isnan(b)) return 0;
    if (a == b) return 0;
    if (std::isnan(a) || std::isnan(b) || std::isinf(a) || std::isinf(b)) return INT64_MAX;
    if (single) {
        const float fa = (float)a, fb = (float)b;
        int32_t ia, ib;
        std::memcpy(&ia, &fa, 4);
        std::memcpy(&ib, &fb, 4);
        if (ia < 0) ia = INT32_MIN - ia;
        if (ib < 0) ib = INT32_MIN - ib;
        return std::llabs((int64_t)ia - (int64_t)ib);
    }
    int64_t ia, ib;
    std::memcpy(&ia, &a, 8);
    std::memcpy(&ib, &b, 8);
    if (ia < 0) ia = INT64_MIN - ia;
    if (ib < 0) ib = INT64_MIN - ib;
    const __int128 d = (__int128)ia - (__int128)ib;
    return d < 0 ? (int64_t)-d : (int64_t)d;
}

template <class T> T *dev_copy(const std::vector<T> &h, hipError_t *e) {
    T *p = nullptr;
    if ((*e = hipMalloc(&p, h.size() * sizeof(T))) != hipSuccess) return nullptr;
    *e = hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    return p;
}

char g_report[8192];

}  // namespace

extern "C" int dc_groups() { return G_COUNT; }

// Text description of the first mismatches of the last dc_run.
extern "C" const char *dc_report() { return g_report; }

extern "C" int dc_run(int g, int n, uint64_t seed, double *max_ulp, int64_t *inexact, int64_t *bad) {
    std::vector<double> X, O(n * (size_t)NO), Oh(n * (size_t)NO);
    std::vector<int> K;
    gen(g, n, seed, X, K);
    // one-cell boundary series: discharge, level, 4-point rating curve, crest
    std::vector<double> QT = {3.0, 3.5}, DQT = {0.01, 0.02}, ZT = {10.8, 10.9}, DZT = {0.002, 0.001};
    std::vector<double> ZW = {10.2, 10.5, 10.9, 11.6}, QW = {0.0, 0.05, 0.2, 1.0}, BF = {10.4};
    std::vector<float> QTf(QT.begin(), QT.end()), DQTf(DQT.begin(), DQT.end()), ZTf(ZT.begin(), ZT.end()),
        DZTf(DZT.begin(), DZT.end()), ZWf(ZW.begin(), ZW.end()), QWf(QW.begin(), QW.end()), BFf(BF.begin(), BF.end());
    std::vector<int32_t> NHQ1 = {NHQ};
    BcArrays H{QT.data(), DQT.data(), ZT.data(), DZT.data(), ZW.data(), QW.data(), BF.data(),
               QTf.data(), DQTf.data(), ZTf.data(), DZTf.data(), ZWf.data(), QWf.data(), BFf.data(), NHQ1.data()};
    hipError_t e = hipSuccess;
    std::vector<void *> owned;
    auto keep = [&](void *p) { owned.push_back(p); return p; };
    BcArrays D;
    D.QT = (double *)keep(dev_copy(QT, &e)); D.DQT = (double *)keep(dev_copy(DQT, &e));
    D.ZT = (double *)keep(dev_copy(ZT, &e)); D.DZT = (double *)keep(dev_copy(DZT, &e));
    D.ZW = (double *)keep(dev_copy(ZW, &e)); D.QW = (double *)keep(dev_copy(QW, &e));
    D.BF = (double *)keep(dev_copy(BF, &e));
    D.QTf = (float *)keep(dev_copy(QTf, &e)); D.DQTf = (float *)keep(dev_copy(DQTf, &e));
    D.ZTf = (float *)keep(dev_copy(ZTf, &e)); D.DZTf = (float *)keep(dev_copy(DZTf, &e));
    D.ZWf = (float *)keep(dev_copy(ZWf, &e)); D.QWf = (float *)keep(dev_copy(QWf, &e));
    D.BFf = (float *)keep(dev_copy(BFf, &e));
    D.NHQ1 = (int32_t *)keep(dev_copy(NHQ1, &e));
    double *dX = (double *)keep(dev_copy(X, &e));
    int *dK = (int *)keep(dev_copy(K, &e));
    double *dO = nullptr;
    if (e == hipSuccess) e = hipMalloc(&dO, O.size() * sizeof(double));
    keep(dO);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_eval, dim3((n + 255) / 256), dim3(256), 0, 0, g, dX, dK, D, dO, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(O.data(), dO, O.size() * sizeof(double), hipMemcpyDeviceToHost);
    for (void *p : owned)
        if (p) (void)hipFree(p);
    if (e != hipSuccess) return (int)e;
    const bool single = (g == G_SLOT32_A || g == G_SLOT32_B || g == G_PAIR32 || g == G_UPDATE32);
    // The cell update divides the friction term by pow(H, 0.33333), whose
    // host build uses a stand-in for the device's reciprocal; a 1-ulp
    // difference there can lose bits to cancellation in the momentum
    // update, so for the update a difference also has to exceed
    // 1e-12 (1 + |H| + |U| + |V|) to count as bad.
    const bool upd = (g == G_UPDATE64 || g == G_UPDATE32 || g == G_UPDATE_BAR);
    double mx = 0;
    int64_t ninex = 0, nbad = 0;
    size_t rep = 0;
    g_report[0] = 0;
    for (int i = 0; i < n; i++) {
        const double *x = &X[(size_t)i * NX];
        eval(g, x, K[i], H, &Oh[(size_t)i * NO]);
        const double floor_abs = upd ? 1e-12 * (1 + fabs(x[0]) + fabs(x[1]) + fabs(x[2])) : 0.0;
        for (int k = 0; k < NO; k++) {
            const double a = O[(size_t)i * NO + k], b = Oh[(size_t)i * NO + k];
            const int64_t d = ulp_diff(a, b, single);
            if (d) ninex++;
            if (d > 4 && !(upd && std::isfinite(a) && std::isfinite(b) && fabs(a - b) <= floor_abs)) {
                if (nbad < 6 && rep < sizeof(g_report) - 600) {
                    rep += snprintf(g_report + rep, sizeof(g_report) - rep,
                                    "case %d kind %d out %d: device %.17g host %.17g | in", i, K[i], k, a, b);
                    for (int q = 0; q < NX; q++)
                        rep += snprintf(g_report + rep, sizeof(g_report) - rep, " %.17g", x[q]);
                    rep += snprintf(g_report + rep, sizeof(g_report) - rep, "\n");
                }
                nbad++;
            }
            if ((double)d > mx) mx = (double)d;
        }
    }
    *max_ulp = mx;
    *inexact = ninex;
    *bad = nbad;
    return 0;
}
