// adjoint.hip -- tangent-linear and adjoint steps of the fp64 timestep on the
// reordered layout (variant A).
//
// One adjoint step maps the adjoint of the state after a step (lam_out) to
// the adjoint of the state before it (lam_in) and accumulates the gradients
// of the per-cell parameters FNC (Manning, 9.81 n^2) and ZBC (bed).  The
// forward step scatters nothing (each cell gathers its slot fluxes), so its
// transpose is a gather too:
//   k_adj_update  per cell: reverse of the cell update -> adjoints of the
//                 cell's flux sums (WH, WU, WV) and of its old state;
//   k_adj_flux    per cell: for each slot, the reverse of this cell's slot
//                 flux (its own-state part) and of the neighbour's twin slot
//                 flux (its neighbour-state part, i.e. this cell), seeded from
//                 the flux-sum adjoints of the cell that owns each slot.
// No atomics: every cell writes only its own adjoints, so results are
// deterministic.
//
// k_tlm_step runs the forward step on dual numbers (dual.hpp): the same
// arithmetic plus one directional derivative, used to check the adjoint.
#include "adjoint.hpp"
#include "dual.hpp"
#include "kernels.hpp"

namespace hydro {

template <> struct StorageOf<ad::Dual> {
    using type = double;
};

namespace {

inline int div_up(int a, int b) { return (a + b - 1) / b; }

__device__ __forceinline__ BcCtx<double> bc_of(const ExtView<double> &m, int cell, int slot, int kt, int jt) {
    BcCtx<double> bc;
    bc.QT = m.QT; bc.DQT = m.DQT; bc.ZT = m.ZT; bc.DZT = m.DZT;
    bc.ZW = m.ZW; bc.QW = m.QW; bc.BF = m.BF; bc.NHQ1 = m.NHQ1;
    bc.CELL = m.CELL; bc.NHQ = m.NHQ; bc.cell = cell; bc.kt = kt; bc.jt = jt;
    bc.SIDE = m.SIDE[slot]; bc.COSF = m.COSF[slot]; bc.SINF = m.SINF[slot];
    return bc;
}

// Boundary context over an fp32 mesh's series, for the adjoint of an fp32
// trajectory: slot_flux_bar reads QT, DQT, ZT, DZT and the slot's SIDE.
struct BcF32 {
    const float *QT, *DQT, *ZT, *DZT;
    int CELL, cell, kt, jt;
    double SIDE;
};

__device__ __forceinline__ BcF32 bc_of(const ExtView<float> &m, int cell, int slot, int kt, int jt) {
    return BcF32{m.QT, m.DQT, m.ZT, m.DZT, m.CELL, cell, kt, jt, (double)m.SIDE[slot]};
}

__device__ __forceinline__ BcCtx<double> bc_none() {
    BcCtx<double> bc;
    bc.QT = bc.DQT = bc.ZT = bc.DZT = bc.ZW = bc.QW = bc.BF = nullptr;
    bc.NHQ1 = nullptr;
    bc.CELL = bc.NHQ = bc.cell = bc.kt = bc.jt = 0;
    bc.SIDE = bc.COSF = bc.SINF = 0;
    return bc;
}

__global__ void k_adj_update(FusedView<double> f, const State<double> *s_in, const double *wsum,
                             const State<double> *lam_out, State<double> *lam_in, double *wbar, double *g_fnc,
                             double *g_zbc) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int C = f.C;
    if (p >= C) return;
    const State<double> s = s_in[p];
    const State<double> l = lam_out[p];
    const UpdateBar b = update_bar(f.S, s.H, s.U, s.V, f.AREA[p], f.FNC[p], wsum[p], wsum[C + p], wsum[2 * C + p],
                                   l.H, l.U, l.V, l.Z);
    wbar[p] = b.WH;
    wbar[C + p] = b.WU;
    wbar[2 * C + p] = b.WV;
    lam_in[p] = State<double>{b.H1, b.U1, b.V1, 0.0};
    g_fnc[p] += b.FNC;
    g_zbc[p] += b.ZBC;
}

__device__ __forceinline__ UpdateBar wbar_of(const double *wbar, int C, int q) {
    UpdateBar u;
    u.WH = wbar[q];
    u.WU = wbar[C + q];
    u.WV = wbar[2 * C + q];
    return u;
}

__global__ void k_adj_flux(FusedView<double> f, const uint8_t *twin, const State<double> *s_in,
                           const double *wbar, State<double> *lam_in, double *g_zbc, int kt, int jt, int *unsupported) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int C = f.C;
    if (p >= C) return;
    const State<double> sp = s_in[p];
    const SideIn me{sp.H, sp.U, sp.V, sp.Z, f.ZBC[p]};
    const UpdateBar wp = wbar_of(wbar, C, p);
    State<double> lam = lam_in[p];
    double gz = 0;
    bool bad = false;
#pragma unroll 1
    for (int j = 0; j < 4; j++) {
        const int e = j * C + p;
        const int kind = f.meta[e] & META_KIND_MASK;
        if (kind == K_PAD) continue;
        const int q = f.nbr[e];
        const bool has = q >= 0;
        SideIn nb{0, 0, 0, 0, 0};
        if (has) {
            const State<double> sq = s_in[q];
            nb = SideIn{sq.H, sq.U, sq.V, sq.Z, f.ZBC[q]};
        }
        const double c = f.COSF[e], s = f.SINF[e], SL = f.SIDE[e];
        const double SLCA = f.SLCOS ? f.SLCOS[e] : SL * c;
        const double SLSA = f.SLSIN ? f.SLSIN[e] : SL * s;
        double Fb[4];
        flux_bar_of_sums(wp, SL, SLCA, SLSA, Fb);
        const BcCtx<double> bc = (kind >= K_B10) ? bc_of(*f.ext, f.orig[p], 4 * f.orig[p] + j, kt, jt) : bc_none();
        SideBar mb, nbb;
        bad |= !slot_flux_bar(kind, f.S, bc, me, has, nb, c, s, Fb, mb, nbb);
        lam.H += mb.H;
        lam.U += mb.U;
        lam.V += mb.V;
        lam.Z += mb.Z;
        gz += mb.ZB;
        if (!has) continue;
        // the neighbour's slot facing this cell: its flux depends on this cell
        const int j2 = twin[e];
        if (j2 > 3) continue;
        const int e2 = j2 * C + q;
        const int kind2 = f.meta[e2] & META_KIND_MASK;
        if (kind2 == K_PAD) continue;
        const double c2 = f.COSF[e2], s2 = f.SINF[e2], SL2 = f.SIDE[e2];
        const double SLCA2 = f.SLCOS ? f.SLCOS[e2] : SL2 * c2;
        const double SLSA2 = f.SLSIN ? f.SLSIN[e2] : SL2 * s2;
        double Fb2[4];
        flux_bar_of_sums(wbar_of(wbar, C, q), SL2, SLCA2, SLSA2, Fb2);
        const BcCtx<double> bc2 =
            (kind2 >= K_B10) ? bc_of(*f.ext, f.orig[q], 4 * f.orig[q] + j2, kt, jt) : bc_none();
        SideBar mb2, nbb2;
        bad |= !slot_flux_bar(kind2, f.S, bc2, nb, true, me, c2, s2, Fb2, mb2, nbb2);
        lam.H += nbb2.H;
        lam.U += nbb2.U;
        lam.V += nbb2.V;
        lam.Z += nbb2.Z;
        gz += nbb2.ZB;
    }
    lam_in[p] = lam;
    g_zbc[p] += gz;
    if (bad) *unsupported = 1;
}

// Adjoint step in one launch.  A workgroup takes kAdjTile consecutive cells
// of the layout (whose tiles keep neighbours close) with eight threads per
// cell: thread (l, j) for slot j of cell l, and twin thread (l, j) for the
// neighbour's slot facing it.
//   1. thread (l, j) reverses its slot's flux (seeded from the flux-sum
//      adjoints of its cell, ub_cur: the reversed cell update of every cell
//      for this step), which gives that flux's adjoint with respect to both
//      cells' states: the own part stays with slot j of this cell; the
//      neighbour's part is the twin part of the neighbour's slot facing this
//      cell, handed over in LDS when the neighbour is among the
//      workgroup's cells;
//   2. for a neighbour outside them, twin thread (l, j) reverses the
//      neighbour's facing slot (seeded from the neighbour's ub_cur) for the
//      twin part;
//   3. thread (l, 0) adds the parts to its cell's update adjoint in slot
//      order, own before twin, the order k_adj_flux uses; both halves of a
//      slot's reversal are the same call's results there and here, so both
//      paths give the same bits;
//   4. with a previous step in the same recorded segment, that thread then
//      reverses its cell's update of the previous step (whose output adjoint
//      it has just formed) into ub_next, so every cell's update is reversed
//      once per step and the next launch's slots read it.
// k_adj_prep fills ub for the last step of a segment.
// Gauge seeds enter where an update's output adjoint is formed: gmap[p] is
// the first gauge index on cell p (-1 none), gnext chains repeated ids in
// list order, and seed[i] is added to the Z adjoint in that order (as
// k_gauge_seed adds it).
constexpr int kAdjTile = 64;
constexpr int kAdjBlock = 8 * kAdjTile;
// one reversed cell update: flux-sum adjoints WH WU WV, old-state adjoints
// H1 U1 V1, and the FNC and ZBC parts
constexpr int kUb = 8;

// The step's scalars in double (an fp32 mesh's are its float values).
template <class R> __device__ __forceinline__ Scalars<double> scalars64(const Scalars<R> &S) {
    return Scalars<double>{(double)S.DT, (double)S.HM1, (double)S.HM2, (double)S.JL};
}

template <class R> __device__ __forceinline__ auto bc_empty_like(const ExtView<R> &) {
    if constexpr (sizeof(R) == 8) {
        return bc_none();
    } else {
        return BcF32{nullptr, nullptr, nullptr, nullptr, 0, 0, 0, 0, 0.0};
    }
}

__device__ __forceinline__ double seeded_lz(double z, int p, const int32_t *gmap, const int32_t *gnext,
                                            const double *seed) {
    if (seed)
        for (int g = gmap[p]; g >= 0; g = gnext[g]) z += seed[g];
    return z;
}

// The reversed update of cell q for the step whose input state and flux sums
// are s_in, wsum, given the adjoint l of its output (seeds not yet added).
template <class R>
__device__ __forceinline__ UpdateBar cell_bar(const FusedView<R> &f, const R *fnc, const State<R> *s_in,
                                              const R *wsum, const State<double> &l, const int32_t *gmap,
                                              const int32_t *gnext, const double *seed, int q) {
    const int C = f.C;
    const State<R> s = s_in[q];
    const double lz = seeded_lz(l.Z, q, gmap, gnext, seed);
    return update_bar(scalars64(f.S), (double)s.H, (double)s.U, (double)s.V, (double)f.AREA[q], (double)fnc[q],
                      (double)wsum[q], (double)wsum[C + q], (double)wsum[2 * C + q], l.H, l.U, l.V, lz);
}

__device__ __forceinline__ void ub_store(double *ub, int C, int p, const UpdateBar &u) {
    ub[p] = u.WH;
    ub[C + p] = u.WU;
    ub[2 * C + p] = u.WV;
    ub[3 * C + p] = u.H1;
    ub[4 * C + p] = u.U1;
    ub[5 * C + p] = u.V1;
    ub[6 * C + p] = u.FNC;
    ub[7 * C + p] = u.ZBC;
}

__device__ __forceinline__ UpdateBar ub_load(const double *ub, int C, int p) {
    UpdateBar u;
    u.WH = ub[p];
    u.WU = ub[C + p];
    u.WV = ub[2 * C + p];
    u.H1 = ub[3 * C + p];
    u.U1 = ub[4 * C + p];
    u.V1 = ub[5 * C + p];
    u.FNC = ub[6 * C + p];
    u.ZBC = ub[7 * C + p];
    return u;
}

// One step's recorded input and flux sums, its gauge seeds, and where the
// reversed updates go; every pointer is scenario 0's (blockIdx.y offsets).
template <class R> struct AdjRec {
    const State<R> *s_in;
    const R *wsum;
    const double *seed;   // null: no gauge seeds
};

template <class R>
__global__ void __launch_bounds__(256) k_adj_prep(FusedView<R> f, AdjRec<R> r, const State<double> *lam_out, double *ub,
                                                  const int32_t *gmap, const int32_t *gnext, int nseed) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int C = f.C;
    if (p >= C) return;
    const size_t so = (size_t)blockIdx.y * C;
    const R *fnc = f.FNC + so;
    const UpdateBar u = cell_bar(f, fnc, r.s_in + so, r.wsum + 3 * so, lam_out[so + p], gmap, gnext,
                                 r.seed ? r.seed + (size_t)blockIdx.y * nseed : nullptr, p);
    ub_store(ub + kUb * so, C, p, u);
}

// R is the trajectory's type: the recorded states, flux sums and the mesh
// statics are read in R and promoted; the adjoint arithmetic, the adjoint
// states and the gradients are double.  blockIdx.y is the scenario of a
// scenario batch (one mesh, S Manning fields and states): its state, flux
// sums, adjoints, gradients, reversed updates and FNC sit blockIdx.y blocks
// of C (3C for the sums, kUb C for the updates) into the arrays, its gauge
// seeds nseed into the seed arrays.
template <class R>
__global__ void __launch_bounds__(kAdjBlock) k_adj_step(FusedView<R> f, const uint8_t *twin,
                                                           const uint8_t *twin_kind, AdjRec<R> cur, AdjRec<R> prev,
                                                           const double *ub_cur, double *ub_next,
                                                           State<double> *lam_in, double *g_fnc, double *g_zbc,
                                                           const int32_t *gmap, const int32_t *gnext, int nseed,
                                                           int kt, int jt, int *unsupported) {
    const size_t so = (size_t)blockIdx.y * f.C;
    const State<R> *s_in = cur.s_in + so;
    const R *fnc = f.FNC + so;
    ub_cur += kUb * so;
    lam_in += so;
    g_fnc += so;
    g_zbc += so;
    // LDS: own and twin parts, [2][4][5][T] doubles
    __shared__ double adj_lds[2 * 4 * 5 * kAdjTile];
    const int C = f.C, T = kAdjTile;
    const int t0 = blockIdx.x * T;
    const bool twin_thread = threadIdx.x >= 4 * T;
    const int jl = twin_thread ? threadIdx.x - 4 * T : threadIdx.x;
    const int j = jl / T, l = jl - j * T;
    const int p = t0 + l;
    const bool valid = p < C;
    auto part = [&](int twin_part, int jj, int comp, int ll) -> double & {
        return adj_lds[((twin_part * 4 + jj) * 5 + comp) * T + ll];
    };
    // every first load is independent of the others and unconditional (a
    // valid slot stands in for the threads without one), so they are all in
    // flight together
    const int e = j * C + (p < C ? p : C - 1);
    const int mk = f.meta[e];
    const int qn = f.nbr[e];
    const int j2r = twin[e];
    const int k2r = twin_kind[e];
    const bool tail = !twin_thread && (j == 0) && valid && ub_next;
    const int kind = valid ? (mk & META_KIND_MASK) : K_PAD;
    const int q = (kind != K_PAD) ? qn : -1;
    // the neighbour's slot facing this one (its twin), when it has a flux
    const bool has_twin = q >= 0 && j2r <= 3 && k2r != K_PAD;
    const bool in_tile = q >= t0 && q < t0 + T;
    if (valid && !twin_thread)
        for (int c = 0; c < 5; c++) part(1, j, c, l) = 0.0;
    __syncthreads();
    bool bad = false;
    const Scalars<double> S = scalars64(f.S);
    // own thread: this slot, with this cell as the own side; twin thread (a
    // neighbour outside the workgroup's cells): the neighbour's facing slot,
    // with the neighbour as the own side
    const bool work = twin_thread ? (kind != K_PAD && has_twin && !in_tile) : kind != K_PAD;
    if (work) {
        {
            const bool tw = twin_thread;
            const State<R> sp = s_in[p];
            const SideIn me{sp.H, sp.U, sp.V, sp.Z, (double)f.ZBC[p]};
            SideIn nb{0, 0, 0, 0, 0};
            if (q >= 0) {
                const State<R> sq = s_in[q];
                nb = SideIn{sq.H, sq.U, sq.V, sq.Z, (double)f.ZBC[q]};
            }
            const int es = tw ? j2r * C + q : e;
            const int ks = tw ? k2r : kind;
            const int oc = tw ? q : p, js = tw ? j2r : j;
            UpdateBar ws;
            ws.WH = ub_cur[oc];
            ws.WU = ub_cur[C + oc];
            ws.WV = ub_cur[2 * C + oc];
            // SIDE*COSF and SIDE*SINF in the trajectory's type, as the
            // forward step forms them
            const R cR = f.COSF[es], snR = f.SINF[es], SLR = f.SIDE[es];
            const double c = cR, sn = snR, SL = SLR;
            const double SLCA = f.SLCOS ? (double)f.SLCOS[es] : (double)(R)(SLR * cR);
            const double SLSA = f.SLSIN ? (double)f.SLSIN[es] : (double)(R)(SLR * snR);
            double Fb[4];
            flux_bar_of_sums(ws, SL, SLCA, SLSA, Fb);
            auto bc = bc_empty_like(*f.ext);
            if (ks >= K_B10) bc = bc_of(*f.ext, f.orig[oc], 4 * f.orig[oc] + js, kt, jt);
            // one call for both passes: the own side is the neighbour on the twin pass
            const SideIn A = tw ? nb : me, B = tw ? me : nb;
            SideBar a_own, a_nb;
            bad |= !slot_flux_bar(ks, S, bc, A, tw || q >= 0, B, c, sn, Fb, a_own, a_nb);
            // pass 0: own part here, the neighbour's part to its twin entry
            // when the neighbour is in the tile; pass 1: this cell's twin part
            const SideBar &mine = tw ? a_nb : a_own;
            const int which = tw ? 1 : 0;
            part(which, j, 0, l) = mine.H;
            part(which, j, 1, l) = mine.U;
            part(which, j, 2, l) = mine.V;
            part(which, j, 3, l) = mine.Z;
            part(which, j, 4, l) = mine.ZB;
            if (!tw && has_twin && in_tile) {
                const int lq = q - t0;
                part(1, j2r, 0, lq) = a_nb.H;
                part(1, j2r, 1, lq) = a_nb.U;
                part(1, j2r, 2, lq) = a_nb.V;
                part(1, j2r, 3, lq) = a_nb.Z;
                part(1, j2r, 4, lq) = a_nb.ZB;
            }
        }
    } else if (valid && !twin_thread) {
        for (int c = 0; c < 5; c++) part(0, j, c, l) = 0.0;
    }
    if (bad) *unsupported = 1;
    __syncthreads();
    if (twin_thread || j != 0 || !valid) return;
    const UpdateBar up = ub_load(ub_cur, C, p);
    State<double> lam{up.H1, up.U1, up.V1, 0.0};
    double gz = 0;
    for (int jj = 0; jj < 4; jj++) {
        for (int kk = 0; kk < 2; kk++) {
            // a padding slot or a missing twin contributes exact zeros, which
            // k_adj_flux skips; adding +0.0 leaves every sum's bits unchanged
            // except -0.0 + 0.0, and lam starts from the update's adjoints
            lam.H += part(kk, jj, 0, l);
            lam.U += part(kk, jj, 1, l);
            lam.V += part(kk, jj, 2, l);
            lam.Z += part(kk, jj, 3, l);
            gz += part(kk, jj, 4, l);
        }
    }
    lam_in[p] = lam;
    g_fnc[p] += up.FNC;
    g_zbc[p] = (g_zbc[p] + up.ZBC) + gz;
    if (tail) {   // the previous step's update, reversed from the adjoint just formed
        const UpdateBar u = cell_bar(f, fnc, prev.s_in + so, prev.wsum + 3 * so, lam, gmap, gnext,
                                     prev.seed ? prev.seed + (size_t)blockIdx.y * nseed : nullptr, p);
        ub_store(ub_next + kUb * so, C, p, u);
    }
}

// Gauge tables of the fused adjoint step: gmap[c] = first index i of the
// gauge list with internal cell c (-1 none), gnext[i] = next index with the
// same cell.  One thread walks the list backwards so the chains run in list
// order.
__global__ void k_gauge_map_clear(int32_t *gmap, int C) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < C) gmap[p] = -1;
}

__global__ void k_gauge_map_build(const int32_t *inv, const int32_t *cells, int n, int32_t *gmap, int32_t *gnext,
                                  int32_t *internal) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int i = n - 1; i >= 0; i--) {
        const int c = inv[cells[i]];
        internal[i] = c;
        gnext[i] = gmap[c];
        gmap[c] = i;
    }
}

// Tangent-linear step: the forward step on dual numbers.
__global__ void k_tlm_step(FusedView<double> f, const State<double> *s_in, const State<double> *ds_in,
                           const double *dfnc, const double *dzbc, State<double> *s_out, State<double> *ds_out,
                           int kt, int jt) {
    using ad::Dual;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int C = f.C;
    if (p >= C) return;
    Scalars<Dual> S;
    S.DT = f.S.DT;
    S.HM1 = f.S.HM1;
    S.HM2 = f.S.HM2;
    S.JL = f.S.JL;
    const State<double> a = s_in[p], da = ds_in[p];
    const Dual H1(a.H, da.H), U1(a.U, da.U), V1(a.V, da.V), Z1(a.Z, da.Z);
    const Dual ZBp(f.ZBC[p], dzbc[p]);
    const Dual ZB1 = ZBp + f.S.HM1;
    Dual WH = 0.0, WU = 0.0, WV = 0.0;
    Dual sd[4];
    for (int j = 0; j < 4; j++) {
        const int e = j * C + p;
        const int kind = f.meta[e] & META_KIND_MASK;
        const double SL = f.SIDE[e];
        sd[j] = SL;
        Flux<Dual> fl;
        bool fired = false;
        if (kind == K_PAD) {
            fl = slot_flux<Dual, 0>(K_MIRROR, S, BcCtx<Dual>{}, H1, U1, V1, Z1, ZB1, ZBp, false, 0.0, 0.0, 0.0,
                                    0.0, 0.0, Dual(0.0), Dual(0.0), Flux<Dual>{0.0, 0.0, 0.0, 0.0}, &fired);
            accumulate<Dual>(WH, WU, WV, fl, Dual(0.0), Dual(0.0), Dual(0.0));
            continue;
        }
        const int q = f.nbr[e];
        const bool has = q >= 0;
        Dual Hn = 0.0, Un = 0.0, Vn = 0.0, Zn = 0.0, ZBn = 0.0;
        if (has) {
            const State<double> b = s_in[q], db = ds_in[q];
            Hn = Dual(b.H, db.H);
            Un = Dual(b.U, db.U);
            Vn = Dual(b.V, db.V);
            Zn = Dual(b.Z, db.Z);
            ZBn = Dual(f.ZBC[q], dzbc[q]);
        }
        const double c = f.COSF[e], s = f.SINF[e];
        BcCtx<Dual> bc{};
        if (kind >= K_B10) {
            const ExtView<double> &m = *f.ext;
            const int oc = f.orig[p], eo = 4 * oc + j;
            bc.QT = m.QT; bc.DQT = m.DQT; bc.ZT = m.ZT; bc.DZT = m.DZT;
            bc.ZW = m.ZW; bc.QW = m.QW; bc.BF = m.BF; bc.NHQ1 = m.NHQ1;
            bc.CELL = m.CELL; bc.NHQ = m.NHQ; bc.cell = oc; bc.kt = kt; bc.jt = jt;
            bc.SIDE = m.SIDE[eo]; bc.COSF = m.COSF[eo]; bc.SINF = m.SINF[eo];
        }
        fl = slot_flux<Dual, 0>(kind, S, bc, H1, U1, V1, Z1, ZB1, ZBp, has, Hn, Un, Vn, Zn, ZBn, Dual(c), Dual(s),
                                Flux<Dual>{0.0, 0.0, 0.0, 0.0}, &fired);
        const double SLCA = f.SLCOS ? f.SLCOS[e] : SL * c;
        const double SLSA = f.SLSIN ? f.SLSIN[e] : SL * s;
        accumulate<Dual>(WH, WU, WV, fl, Dual(SL), Dual(SLCA), Dual(SLSA));
    }
    const Dual FNC(f.FNC[p], dfnc[p]);
    const CellOut<Dual> o = update_math<Dual, 0>(S, H1, U1, V1, ZBp, Dual(f.AREA[p]), FNC, WH, WU, WV, sd[0], sd[1],
                                                 sd[2], sd[3]);
    s_out[p] = State<double>{o.H.v, o.U.v, o.V.v, o.Z.v};
    ds_out[p] = State<double>{o.H.d, o.U.d, o.V.d, o.Z.d};
}

// gauges: value of Z (or add a seed into lam.Z) at ORIGINAL cell ids
// Z of the gauges of S scenario states (st [S][C]) to out [S][n].
template <class R>
__global__ void k_gauge_gather_scen(const State<R> *st, int C, const int32_t *inv, const int32_t *cells, int n,
                                    double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[(size_t)blockIdx.y * n + i] = st[(size_t)blockIdx.y * C + inv[cells[i]]].Z;
}

// External-order double values [S][C] -> internal-order R values [S][C].
template <class R>
__global__ void k_scen_to_internal(const double *in, const int32_t *orig, R *out, int C) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < C) out[(size_t)blockIdx.y * C + p] = (R)in[(size_t)blockIdx.y * C + orig[p]];
}

template <class R>
__global__ void k_gauge_gather(const State<R> *st, const int32_t *inv, const int32_t *cells, int n, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = st[inv[cells[i]]].Z;
}

__global__ void k_gauge_seed(State<double> *lam, const int32_t *inv, const int32_t *cells, int n, const double *seed) {
    // one thread walks the list so repeated gauge ids accumulate deterministically
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int i = 0; i < n; i++) lam[inv[cells[i]]].Z += seed[i];
}

__global__ void k_to_external(const double *in, const int32_t *orig, double *out, int C) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < C) out[orig[p]] = in[p];
}

__global__ void k_to_internal(const double *in, const int32_t *orig, double *out, int C) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < C) out[p] = in[orig[p]];
}

__global__ void k_states_to_external(const State<double> *in, const int32_t *orig, double *H, double *U, double *V,
                                     double *Z, int C) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= C) return;
    const State<double> s = in[p];
    const int o = orig[p];
    H[o] = s.H;
    U[o] = s.U;
    V[o] = s.V;
    Z[o] = s.Z;
}

__global__ void k_states_to_internal(const double *H, const double *U, const double *V, const double *Z,
                                     const int32_t *orig, State<double> *out, int C) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= C) return;
    const int o = orig[p];
    out[p] = State<double>{H[o], U[o], V[o], Z[o]};
}

}  // namespace

hipError_t launch_adjoint_step(const FusedView<double> &f, const uint8_t *twin, const State<double> *s_in,
                               const double *wsum, const State<double> *lam_out, State<double> *lam_in, double *wbar,
                               double *g_fnc, double *g_zbc, int kt, int jt, int *unsupported, hipStream_t s) {
    if (f.C == 0) return hipSuccess;
    const int B = 256;
    hipLaunchKernelGGL(k_adj_update, dim3(div_up(f.C, B)), dim3(B), 0, s, f, s_in, wsum, lam_out, lam_in, wbar, g_fnc,
                       g_zbc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_adj_flux, dim3(div_up(f.C, B)), dim3(B), 0, s, f, twin, s_in, wbar, lam_in, g_zbc, kt, jt,
                       unsupported);
    return hipGetLastError();
}

template <class R>
hipError_t launch_adjoint_prep(const FusedView<R> &f, const State<R> *s_in, const R *wsum, const double *seed,
                               const State<double> *lam_out, double *ub, const int32_t *gmap, const int32_t *gnext,
                               int nseed, int nscen, hipStream_t s) {
    if (f.C == 0 || nscen <= 0) return hipSuccess;
    const AdjRec<R> r{s_in, wsum, seed};
    hipLaunchKernelGGL((k_adj_prep<R>), dim3(div_up(f.C, 256), nscen), dim3(256), 0, s, f, r, lam_out, ub, gmap, gnext,
                       nseed);
    return hipGetLastError();
}

template <class R>
hipError_t launch_adjoint_step_fused(const FusedView<R> &f, const uint8_t *twin, const uint8_t *twin_kind,
                                     const State<R> *s_in, const R *wsum,
                                     const State<R> *s_prev, const R *wsum_prev, const double *seed_prev,
                                     const double *ub_cur, double *ub_next, State<double> *lam_in, double *g_fnc,
                                     double *g_zbc, const int32_t *gmap, const int32_t *gnext, int nseed, int nscen,
                                     int kt, int jt, int *unsupported, hipStream_t s) {
    if (f.C == 0 || nscen <= 0) return hipSuccess;
    const AdjRec<R> cur{s_in, wsum, nullptr}, prev{s_prev, wsum_prev, seed_prev};
    hipLaunchKernelGGL((k_adj_step<R>), dim3(div_up(f.C, kAdjTile), nscen), dim3(kAdjBlock), 0, s, f, twin, twin_kind,
                       cur, prev, ub_cur, ub_next, lam_in, g_fnc, g_zbc, gmap, gnext, nseed, kt, jt, unsupported);
    return hipGetLastError();
}

#define HYDRO_ADJ_INST(R)                                                                                        \
    template hipError_t launch_adjoint_prep<R>(const FusedView<R> &, const State<R> *, const R *, const double *, \
                                               const State<double> *, double *, const int32_t *, const int32_t *, \
                                               int, int, hipStream_t);                                          \
    template hipError_t launch_adjoint_step_fused<R>(const FusedView<R> &, const uint8_t *, const uint8_t *,      \
                                                     const State<R> *,                                           \
                                                     const R *, const State<R> *, const R *, const double *,      \
                                                     const double *, double *, State<double> *, double *, double *, \
                                                     const int32_t *, const int32_t *, int, int, int, int, int *, \
                                                     hipStream_t);
HYDRO_ADJ_INST(double)
HYDRO_ADJ_INST(float)

hipError_t launch_gauge_map(const int32_t *inv, const int32_t *cells, int n, int C, int32_t *gmap, int32_t *gnext,
                            int32_t *internal, hipStream_t s) {
    if (C > 0) hipLaunchKernelGGL(k_gauge_map_clear, dim3(div_up(C, 256)), dim3(256), 0, s, gmap, C);
    if (n > 0) hipLaunchKernelGGL(k_gauge_map_build, dim3(1), dim3(64), 0, s, inv, cells, n, gmap, gnext, internal);
    return hipGetLastError();
}

hipError_t launch_tlm_step(const FusedView<double> &f, const State<double> *s_in, const State<double> *ds_in,
                           const double *dfnc, const double *dzbc, State<double> *s_out, State<double> *ds_out, int kt,
                           int jt, hipStream_t s) {
    if (f.C == 0) return hipSuccess;
    hipLaunchKernelGGL(k_tlm_step, dim3(div_up(f.C, 128)), dim3(128), 0, s, f, s_in, ds_in, dfnc, dzbc, s_out, ds_out,
                       kt, jt);
    return hipGetLastError();
}

template <class R>
hipError_t launch_gauge_gather(const State<R> *st, const int32_t *inv, const int32_t *cells, int n, double *out,
                               hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL((k_gauge_gather<R>), dim3(div_up(n, 256)), dim3(256), 0, s, st, inv, cells, n, out);
    return hipGetLastError();
}
template hipError_t launch_gauge_gather<double>(const State<double> *, const int32_t *, const int32_t *, int, double *,
                                                hipStream_t);

template <class R>
hipError_t launch_gauge_gather_scen(const State<R> *st, int C, int nscen, const int32_t *inv, const int32_t *cells,
                                    int n, double *out, hipStream_t s) {
    if (n == 0 || nscen <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_gauge_gather_scen<R>), dim3(div_up(n, 256), nscen), dim3(256), 0, s, st, C, inv, cells, n,
                       out);
    return hipGetLastError();
}
template hipError_t launch_gauge_gather_scen<double>(const State<double> *, int, int, const int32_t *,
                                                     const int32_t *, int, double *, hipStream_t);
template hipError_t launch_gauge_gather_scen<float>(const State<float> *, int, int, const int32_t *, const int32_t *,
                                                    int, double *, hipStream_t);

template <class R>
hipError_t launch_scen_to_internal(const double *in, const int32_t *orig, R *out, int C, int nscen, hipStream_t s) {
    if (C == 0 || nscen <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_scen_to_internal<R>), dim3(div_up(C, 256), nscen), dim3(256), 0, s, in, orig, out, C);
    return hipGetLastError();
}
template hipError_t launch_scen_to_internal<double>(const double *, const int32_t *, double *, int, int, hipStream_t);
template hipError_t launch_scen_to_internal<float>(const double *, const int32_t *, float *, int, int, hipStream_t);
template hipError_t launch_gauge_gather<float>(const State<float> *, const int32_t *, const int32_t *, int, double *,
                                               hipStream_t);

hipError_t launch_gauge_seed(State<double> *lam, const int32_t *inv, const int32_t *cells, int n, const double *seed,
                             hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gauge_seed, dim3(1), dim3(64), 0, s, lam, inv, cells, n, seed);
    return hipGetLastError();
}

hipError_t launch_cells_map(const double *in, const int32_t *orig, double *out, int C, int to_external, hipStream_t s) {
    if (C == 0) return hipSuccess;
    if (to_external)
        hipLaunchKernelGGL(k_to_external, dim3(div_up(C, 256)), dim3(256), 0, s, in, orig, out, C);
    else
        hipLaunchKernelGGL(k_to_internal, dim3(div_up(C, 256)), dim3(256), 0, s, in, orig, out, C);
    return hipGetLastError();
}

hipError_t launch_states_map(State<double> *st, const int32_t *orig, double *H, double *U, double *V, double *Z, int C,
                             int to_external, hipStream_t s) {
    if (C == 0) return hipSuccess;
    if (to_external)
        hipLaunchKernelGGL(k_states_to_external, dim3(div_up(C, 256)), dim3(256), 0, s, st, orig, H, U, V, Z, C);
    else
        hipLaunchKernelGGL(k_states_to_internal, dim3(div_up(C, 256)), dim3(256), 0, s, H, U, V, Z, orig, st, C);
    return hipGetLastError();
}

}  // namespace hydro
