// kernels.hip -- gfx950 kernels of the hydro-cal timestep.
//
//  k_calculate_flux / k_update_cell : the reference's two-kernel contract on
//      the caller's layout (CalculateFluxKernel F:47-202, UpdateCellKernel
//      F:717-784 of cuda_native_impl/hydro-cal-src/src/functors.cu).
//  k_fused_step : one thread per cell of the reordered layout computes its
//      four slot fluxes in registers and applies the update, reading state
//      record `in` and writing record `out` (ping-pong replaces the
//      reference's in-place update after a separate flux pass).
//  k_persistent : the fused step for many steps in one launch, with a
//      grid-wide barrier between steps (the design of
//      cuda_impl/F2_driver.cu:34-44, written natively).
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

#include "ext_step.hpp"
#include "kernels.hpp"

namespace hydro {

namespace {

// Tile index of this workgroup.  Workgroups are dispatched round-robin over
// the 8 XCDs of MI355X; with remap on, XCD x gets the contiguous tile range
// [x*q + min(x,r), ...) so neighbouring cells share that XCD's L2.
__device__ __forceinline__ int tile_of(int wg, int nwg, int remap) {
    if (!remap) return wg;
    const int q = nwg >> 3, r = nwg & 7;
    const int x = wg & 7, k = wg >> 3;
    return x * q + (x < r ? x : r) + k;
}

// ---------------------------------------------------------------------------
// Two-kernel path on the caller's layout.

template <class R, int VAR>
__global__ void __launch_bounds__(kMaxBlock) k_calculate_flux(ExtView<R> m, int kt, int jt) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < 4 * m.CELL) ext_flux_slot<R, VAR>(m, idx, kt, jt);
}

template <class R, int VAR>
__global__ void __launch_bounds__(kMaxBlock) k_update_cell(ExtView<R> m) {
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos < m.CELL) ext_update_cell<R, VAR>(m, pos);
}

// ---------------------------------------------------------------------------
// Fused step on the reordered layout.

// Gauge levels of a step's input state: cell p's value to every gauge on p.
template <class R>
__device__ __forceinline__ void gauge_write(const FusedView<R> &f, int p, R z, int scen) {
    double *out = f.gauge_out + (size_t)scen * f.gauge_n;
    for (int g = f.gauge_map[p]; g >= 0; g = f.gauge_next[g]) out[g] = (double)z;
}

// RARE = false: the mesh has only interior, KLAS 10/1/4 and padding slots
// (checked on the host), so the other boundary handlers are not compiled in.
template <class R, int VAR, bool RARE>
__device__ __forceinline__ void fused_cell(const FusedView<R> &f, const State<R> *__restrict__ Sin,
                                           State<R> *__restrict__ Sout, int p, int kt, int jt) {
    const int C = f.C;
    const Scalars<R> S = f.S;
    const State<R> me = Sin[p];
    const R ZBp = f.ZBC[p];
    const R ZB1 = f.ZB1 ? f.ZB1[p] : (R)(ZBp + S.HM1);
    R WH = 0.0, WU = 0.0, WV = 0.0;
    R sd[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int e = j * C + p;
        const int mt = f.meta[e];
        const int kind = mt & META_KIND_MASK;
        const R SIDE = f.SIDE[e];
        sd[j] = SIDE;
        Flux<R> fl;
        if (kind == K_PAD) {
            if (f.materialize) {
                bool fired = false;
                fl = slot_flux<R, VAR, RARE>(K_MIRROR, S, empty_bc<R>(), me.H, me.U, me.V, me.Z, ZB1, ZBp,
                                       false, 0, 0, 0, 0, 0, (R)0, (R)0, Flux<R>{0, 0, 0, 0}, &fired);
            } else {
                fl = padding_flux_signs<R>(S, me.H, me.U, me.V, me.Z, ZB1, ZBp);
            }
            accumulate<R>(WH, WU, WV, fl, (R)0, (R)0, (R)0);
        } else {
            const int nb = f.nbr[e];
            const bool has = nb >= 0;
            State<R> o = {0, 0, 0, 0};
            R ZBn = 0;
            if (has) {
                o = Sin[nb];
                ZBn = f.ZBC[nb];
            }
            const R c = f.COSF[e], s = f.SINF[e];
            Flux<R> prev = {0, 0, 0, 0};
            BcCtx<R> bc = empty_bc<R>();
            if (kind >= K_B10) {
                const ExtView<R> &m = *f.ext;
                const int eo = 4 * f.orig[p] + j;
                bc = make_bc<R>(m, f.orig[p], eo, kt, jt);
                if (RARE && kind == K_BUNK) prev = {m.F0[eo], m.F1[eo], m.F2[eo], m.F3[eo]};
            }
            bool fired = false;
            fl = slot_flux<R, VAR, RARE>(kind, S, bc, me.H, me.U, me.V, me.Z, ZB1, ZBp, has, o.H, o.U, o.V,
                                   o.Z, ZBn, c, s, prev, &fired);
            if (RARE && fired) {
                f.meta[e] = (mt & META_PRIMARY_BIT) ? (uint8_t)(K_PRIMARY | META_PRIMARY_BIT)
                                                    : (uint8_t)K_MIRROR;
                f.ext->KLAS[4 * f.orig[p] + j] = 0;
                if (f.ext->KLASr) f.ext->KLASr[4 * f.orig[p] + j] = 0;
            }
            const R SLCA = f.SLCOS ? f.SLCOS[e] : SIDE * c;
            const R SLSA = f.SLSIN ? f.SLSIN[e] : SIDE * s;
            accumulate<R>(WH, WU, WV, fl, SIDE, SLCA, SLSA);
        }
        // A slot whose code has no handler keeps FLUX0/1/3 from its previous
        // step (and the supercritical branch writes all four), so its flux
        // always goes to the caller's arrays.
        if (f.materialize || (RARE && kind == K_BUNK)) {
            const ExtView<R> &m = *f.ext;
            const int eo = 4 * f.orig[p] + j;
            m.F0[eo] = fl.f0; m.F1[eo] = fl.f1; m.F2[eo] = fl.f2; m.F3[eo] = fl.f3;
        }
    }
    if (f.rec_state) f.rec_state[p] = me;
    if (f.rec_wsum) {
        f.rec_wsum[p] = WH;
        f.rec_wsum[C + p] = WU;
        f.rec_wsum[2 * C + p] = WV;
    }
    if (f.gauge_out) gauge_write(f, p, me.Z, 0);
    const CellOut<R> o = update_math<R, VAR>(S, me.H, me.U, me.V, ZBp, f.AREA[p], f.FNC[p], WH, WU, WV,
                                             sd[0], sd[1], sd[2], sd[3]);
    Sout[p] = State<R>{o.H, o.U, o.V, o.Z};   // W = speed(U, V) is formed on the way out
}

// ---------------------------------------------------------------------------
// TILES scheme: one workgroup per tile of blockDim.x consecutive cells.
//  1. stage the tile's cell states and beds in LDS (and prefetch the tile's
//     task entries and the update's statics into registers);
//  2. run the tile's tasks: an edge joining two cells of the tile is solved
//     once for both slots (pair_flux), any other slot alone (slot_flux);
//     each slot's three flux-sum terms go to LDS;
//  3. each thread sums its cell's four slots in slot order (as the
//     reference's UpdateCellKernel does, F:732-742) and updates the cell.
// Phase 3 reads only LDS terms and registers, so a following tile's phase 1
// cannot overwrite anything still in use once phase 2's barrier is passed.

template <class R> struct TileLds {
    State<R> *st;   // [B]
    R *zb;          // [B] bed ZBC
    R *zb1;         // [B] ZB1
    R *c;           // [3][4][B]: WH, WU, WV terms of slot j of local cell l
};

template <class R> __device__ __forceinline__ TileLds<R> tile_lds(int B) {
    extern __shared__ __align__(32) unsigned char hydro_tile_smem[];
    TileLds<R> L;
    L.st = reinterpret_cast<State<R> *>(hydro_tile_smem);
    L.zb = reinterpret_cast<R *>(hydro_tile_smem + sizeof(State<R>) * B);
    L.zb1 = L.zb + B;
    L.c = L.zb1 + B;
    return L;
}

// Writes the slot's terms to LDS; the flux itself goes to the caller's
// FLUX0..3 when materialising, or when `keep` (a slot whose code has no
// handler reads its previous FLUX0/1/3 next step).
// Writes the slot's terms to LDS; the flux itself goes to the caller's
// FLUX0..3 when materialising, or when `keep` (a slot whose code has no
// handler reads its previous FLUX0/1/3 next step).  SL* are SIDE*COSF and
// SIDE*SINF unless the caller's SLCOS/SLSIN differ from those products.
template <class R>
__device__ __forceinline__ void tile_store(const FusedView<R> &f, const TileLds<R> &L, int B, int l, int j, int p,
                                           const Flux<R> &fl, R SIDE, R c, R s, bool keep = false) {
    R SLCA = SIDE * c, SLSA = SIDE * s;
    if (f.SLCOS) {
        const int e = j * f.C + p;
        SLCA = f.SLCOS[e];
        SLSA = f.SLSIN[e];
    }
    const Contrib<R> t = contrib<R>(fl, SIDE, SLCA, SLSA);
    L.c[j * B + l] = t.h;
    L.c[(4 + j) * B + l] = t.u;
    L.c[(8 + j) * B + l] = t.v;
    if (f.materialize || keep) {
        const ExtView<R> &m = *f.ext;
        const int eo = 4 * f.orig[p] + j;
        m.F0[eo] = fl.f0; m.F1[eo] = fl.f1; m.F2[eo] = fl.f2; m.F3[eo] = fl.f3;
    }
}

// State records as the persistent TILES kernel moves them between
// workgroups: buffer loads and stores with the sc1 cache policy, i.e.
// coherent across the XCDs' L2 caches (what device-scope relaxed atomics
// compile to, but 16 bytes per instruction), so a step's output is visible
// to a neighbouring tile on another XCD once its flag is, with no
// cache-wide write-back or invalidate.  Other launches use plain accesses.
typedef int hydro_i4 __attribute__((ext_vector_type(4)));
constexpr int kBufSC1 = 16;                  // buffer cache-policy operand: sc1
constexpr int kBufWord3 = 0x00020000;        // gfx9 buffer descriptor word 3 (32-bit data format)

template <class R> __device__ __forceinline__ __amdgpu_buffer_rsrc_t state_rsrc(const State<R> *S, int C) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<State<R> *>(S), 0, C * (int)sizeof(State<R>), kBufWord3);
}

template <bool COH, class R> __device__ __forceinline__ State<R> ld_state(const State<R> *S, int i, int C) {
    if (!COH) return S[i];
    const __amdgpu_buffer_rsrc_t r = state_rsrc(S, C);
    const int off = i * (int)sizeof(State<R>);
    if constexpr (sizeof(R) == 8) {
        const hydro_i4 a = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, kBufSC1);
        const hydro_i4 b = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16, 0, kBufSC1);
        struct P {
            hydro_i4 a, b;
        } pk{a, b};
        return __builtin_bit_cast(State<R>, pk);
    } else {
        return __builtin_bit_cast(State<R>, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, kBufSC1));
    }
}

template <bool COH, class R> __device__ __forceinline__ void st_state(State<R> *S, int i, int C, const State<R> &v) {
    if (!COH) {
        S[i] = v;
        return;
    }
    const __amdgpu_buffer_rsrc_t r = state_rsrc(S, C);
    const int off = i * (int)sizeof(State<R>);
    if constexpr (sizeof(R) == 8) {
        struct P {
            hydro_i4 a, b;
        };
        const P pk = __builtin_bit_cast(P, v);
        __builtin_amdgcn_raw_buffer_store_b128(pk.a, r, off, 0, kBufSC1);
        __builtin_amdgcn_raw_buffer_store_b128(pk.b, r, off + 16, 0, kBufSC1);
    } else {
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(hydro_i4, v), r, off, 0, kBufSC1);
    }
}

// One task entry as loaded (all of it in one coalesced round trip), and
// for a one-slot task the neighbour's state and bed.
template <class R> struct TaskIn {
    uint32_t w;
    int nb;
    R c, s, sa, sb;
    State<R> on;
    R zbn;
};

// Neighbour of a one-slot task, fetched as soon as the entry is in (cell
// `own` stands in when there is none, so the load is unconditional).
template <bool COH, class R>
__device__ __forceinline__ void task_neighbour(const FusedView<R> &f, const State<R> *__restrict__ Sin, TaskIn<R> &t,
                                               int own) {
    const bool single = t.w != TASK_NONE && !(t.w & TASK_PAIR);
    const int ni = (single && t.nb >= 0) ? t.nb : own;
    t.on = ld_state<COH>(Sin, ni, f.C);
    t.zbn = f.ZBC[ni];
}

template <class R> __device__ __forceinline__ TaskIn<R> none_task() {
    TaskIn<R> t;
    t.w = TASK_NONE;
    t.nb = -1;
    t.c = t.s = t.sa = t.sb = 0;
    return t;
}

template <class R> __device__ __forceinline__ TaskIn<R> task_load(const FusedView<R> &f, int k) {
    TaskIn<R> t;
    t.w = f.task_w[k];
    t.nb = f.task_nb[k];
    t.c = f.task_cs[2 * k];
    t.s = f.task_cs[2 * k + 1];
    t.sa = f.task_sd[k];
    t.sb = t.sa;
    return t;
}

template <class R, int VAR, bool RARE>
__device__ __forceinline__ void task_run(const FusedView<R> &f, const TileLds<R> &L, const TaskIn<R> &t,
                                         const State<R> *__restrict__ Sin, int B, int p0, int kt, int jt) {
    const Scalars<R> S = f.S;
    const int C = f.C;
    const int la = task_la(t.w), ja = task_ja(t.w);
    const int pa = p0 + la;
    const State<R> A = L.st[la];
    const R zbA = L.zb[la];
    const R zb1A = f.ZB1 ? L.zb1[la] : (R)(zbA + S.HM1);
    if (t.w & TASK_PAIR) {
        const int lb = task_lb(t.w), jb = task_jb(t.w);
        const int pb = p0 + lb;
        const State<R> M = L.st[lb];
        const R zbM = L.zb[lb];
        const R zb1M = f.ZB1 ? L.zb1[lb] : (R)(zbM + S.HM1);
        Flux<R> fA, fM;
        pair_flux<R>(S, A.H, A.U, A.V, A.Z, zb1A, zbA, M.H, M.U, M.V, M.Z, zb1M, zbM, t.c, t.s, &fA, &fM);
        tile_store<R>(f, L, B, la, ja, pa, fA, t.sa, t.c, t.s);
        // the mirror slot's normal is exactly the negated one (a pairing condition)
        tile_store<R>(f, L, B, lb, jb, pb, fM, t.sb, -t.c, -t.s);
        return;
    }
    int kind = task_kind(t.w);
    int mt = 0;
    const int ea = ja * C + pa;
    if (RARE && kind == K_B7) {   // a breach may have opened: its kind is in meta
        mt = f.meta[ea];
        kind = mt & META_KIND_MASK;
    }
    const bool has = t.nb >= 0;
    State<R> o = {0, 0, 0, 0};
    R ZBn = 0;
    if (has) {
        o = t.on;
        ZBn = t.zbn;
    }
    Flux<R> prev = {0, 0, 0, 0};
    BcCtx<R> bc = empty_bc<R>();
    if (kind >= K_B10) {
        const ExtView<R> &m = *f.ext;
        const int eo = 4 * f.orig[pa] + ja;
        bc = make_bc<R>(m, f.orig[pa], eo, kt, jt);
        if (RARE && kind == K_BUNK) prev = {m.F0[eo], m.F1[eo], m.F2[eo], m.F3[eo]};
    }
    bool fired = false;
    const Flux<R> fl = slot_flux<R, VAR, RARE>(kind, S, bc, A.H, A.U, A.V, A.Z, zb1A, zbA, has, o.H, o.U, o.V,
                                               o.Z, ZBn, t.c, t.s, prev, &fired);
    if (RARE && fired) {
        f.meta[ea] = (mt & META_PRIMARY_BIT) ? (uint8_t)(K_PRIMARY | META_PRIMARY_BIT) : (uint8_t)K_MIRROR;
        f.ext->KLAS[4 * f.orig[pa] + ja] = 0;
        if (f.ext->KLASr) f.ext->KLASr[4 * f.orig[pa] + ja] = 0;
    }
    tile_store<R>(f, L, B, la, ja, pa, fl, t.sa, t.c, t.s, RARE && kind == K_BUNK);
}

// Launches that start a step only once the previous one has ended have
// nothing to wait for inside a tile.
struct NoWait {
    __device__ bool operator()() const { return true; }
};

// SCEN: one of several scenarios of the mesh (k_fused_step_scen): Sin and
// Sout are the scenario's, and its FNC and recording buffers sit `scen`
// blocks of C (3C for the flux sums) into the arrays.
// wait(): called by every thread once the tile's own loads are issued and
// before any neighbour state is read (persistent launches wait there for the
// neighbouring tiles' previous step); false ends the tile at once, and
// tile_cells returns false.
template <class R, int VAR, bool RARE, bool SCEN>
__device__ __forceinline__ State<R> tile_finish(const FusedView<R> &f, const TileLds<R> &L, int B, int i, int p,
                                                const State<R> &me, R zbp, R area, R fnc, uint32_t pad,
                                                const R (&sd)[4], int scen);

template <class R, int VAR, bool RARE, bool COH = false, bool SCEN = false, class Wait = NoWait>
__device__ __forceinline__ bool tile_cells(const FusedView<R> &f, const State<R> *__restrict__ Sin,
                                           State<R> *__restrict__ Sout, int t, int kt, int jt, int scen = 0,
                                           Wait wait = Wait()) {
    const int B = f.block, C = f.C;
    // In the persistent kernels (the ones that wait) the thread index passes
    // through an empty asm, so nothing derived from it (LDS and array
    // offsets) is hoisted out of their step and tile loops, where it would
    // hold registers across the whole body (they stay within 128 VGPRs only
    // without such hoisting).  Single-step kernels keep the plain index.
    int i = threadIdx.x;
    if constexpr (!std::is_same<Wait, NoWait>::value) asm volatile("" : "+v"(i));
    const int p0 = t * f.tile, p = p0 + i;
    const bool live = i < f.tile && p < C;
    const TileLds<R> L = tile_lds<R>(B);
    // Phase 1: the cell's state and statics, and the tile's first two task
    // entries per thread, all independent loads in flight together.
    // Every load is issued before the first use of any of them, so they are
    // all in flight together (the compiler orders waits by first use).
    State<R> me = {0, 0, 0, 0};
    R zbp = 0, z1 = 0, area = 0, fnc = 0;
    R sd[4] = {0, 0, 0, 0};
    uint32_t pad = 0;
    const int pc = live ? p : C - 1;   // a valid cell for the unused lanes of the last tile
    me = ld_state<COH>(Sin, pc, C);
    zbp = f.ZBC[pc];
    // ZB1 is one unconditional load (ZBC again when ZB1 is implied): a
    // conditional load here would make its wait a wait for everything
    z1 = (f.ZB1 ? f.ZB1 : f.ZBC)[pc];
    area = (VAR == 0) ? f.DTA[pc] : f.AREA[pc];   // variant A: DT / AREA itself
    fnc = SCEN ? f.FNC[(size_t)scen * C + pc] : f.FNC[pc];
    pad = f.cell_pad[pc];
    if (VAR == 1) {
#pragma unroll
        for (int j = 0; j < 4; j++) sd[j] = f.SIDE[j * C + pc];
    }
    // a wavefront reads the entries of its first two positions when the tile
    // uses them (the test is uniform over the wavefront: positions are 64
    // entries); the positions a tile leaves unused are not read
    const int k0 = t * f.task_stride, kn = k0 + 64 * f.task_n[t];
    const int kw = k0 + (i & ~63);
    TaskIn<R> t0 = none_task<R>(), t1 = none_task<R>();
    if (kw < kn) t0 = task_load<R>(f, k0 + i);
    if (kw + B < kn) t1 = task_load<R>(f, k0 + i + B);
    if (!wait()) return false;
    task_neighbour<COH>(f, Sin, t0, pc);
    task_neighbour<COH>(f, Sin, t1, pc);
    if (live) {
        L.st[i] = me;
        L.zb[i] = zbp;
        if (f.ZB1) L.zb1[i] = z1;   // else ZB1 == ZBC + HM1, computed where used
    }
    __syncthreads();
    // Phase 2: the tasks; each writes its slots' terms to LDS.
    if (t0.w != TASK_NONE) task_run<R, VAR, RARE>(f, L, t0, Sin, B, p0, kt, jt);
    if (t1.w != TASK_NONE) task_run<R, VAR, RARE>(f, L, t1, Sin, B, p0, kt, jt);
    for (int k = k0 + i + 2 * B; k < kn; k += B) {
        TaskIn<R> tk = task_load<R>(f, k);
        task_neighbour<COH>(f, Sin, tk, pc);
        if (tk.w != TASK_NONE) task_run<R, VAR, RARE>(f, L, tk, Sin, B, p0, kt, jt);
    }
    __syncthreads();
    if (!live) return true;
    st_state<COH>(Sout, p, C, tile_finish<R, VAR, RARE, SCEN>(f, L, B, i, p, me, zbp, area, fnc, pad, sd, scen));
    return true;
}

// Phase 3 of a tile step for the thread of live cell p (local index i): the
// cell's four slots in slot order from the LDS terms, then its update; the
// new state is returned (W = speed(U, V) is formed on the way out).
template <class R, int VAR, bool RARE, bool SCEN>
__device__ __forceinline__ State<R> tile_finish(const FusedView<R> &f, const TileLds<R> &L, int B, int i, int p,
                                                const State<R> &me, R zbp, R area, R fnc, uint32_t pad,
                                                const R (&sd)[4], int scen) {
    const Scalars<R> S = f.S;
    const int C = f.C;
    const R zb1 = f.ZB1 ? L.zb1[i] : (R)(zbp + S.HM1);
    R WH = 0.0, WU = 0.0, WV = 0.0;
    if (f.materialize) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if ((pad >> j) & 1) {
                bool fired = false;
                const Flux<R> fl = slot_flux<R, VAR, RARE>(K_MIRROR, S, empty_bc<R>(), me.H, me.U, me.V, me.Z, zb1,
                                                           zbp, false, 0, 0, 0, 0, 0, (R)0, (R)0,
                                                           Flux<R>{0, 0, 0, 0}, &fired);
                const ExtView<R> &m = *f.ext;
                const int eo = 4 * f.orig[p] + j;
                m.F0[eo] = fl.f0; m.F1[eo] = fl.f1; m.F2[eo] = fl.f2; m.F3[eo] = fl.f3;
                accumulate<R>(WH, WU, WV, fl, (R)0, (R)0, (R)0);
            } else {
                WH += L.c[j * B + i];
                WU += L.c[(4 + j) * B + i];
                WV += L.c[(8 + j) * B + i];
            }
        }
    } else {
        // A padding slot adds SIDE * flux = +-0 to each sum.  Adding a zero
        // leaves every nonzero partial sum as it is, and a zero partial sum
        // followed by a nonzero term gives that term, so the padding slots
        // can change a sum only when it comes out exactly zero, and then only
        // its sign.  The sums are formed without them, and only a cell with
        // a zero sum forms them again in slot order with the padding slots'
        // flux signs (the same additions in the same order as the reference).
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (!((pad >> j) & 1)) {
                WH += L.c[j * B + i];
                WU += L.c[(4 + j) * B + i];
                WV += L.c[(8 + j) * B + i];
            }
        if ((pad & 15u) && (WH == (R)0 || WU == (R)0 || WV == (R)0)) {
            const Flux<R> fl = padding_flux_signs<R>(S, me.H, me.U, me.V, me.Z, zb1, zbp);
            WH = 0.0;
            WU = 0.0;
            WV = 0.0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if ((pad >> j) & 1) {
                    accumulate<R>(WH, WU, WV, fl, (R)0, (R)0, (R)0);
                } else {
                    WH += L.c[j * B + i];
                    WU += L.c[(4 + j) * B + i];
                    WV += L.c[(8 + j) * B + i];
                }
            }
        }
    }
    if (f.rec_state) f.rec_state[(SCEN ? (size_t)scen * C : 0) + p] = me;
    if (f.gauge_out) gauge_write(f, p, me.Z, scen);
    if (f.rec_wsum) {
        const size_t so = SCEN ? (size_t)scen * C : 0;
        f.rec_wsum[3 * so + p] = WH;
        f.rec_wsum[3 * so + C + p] = WU;
        f.rec_wsum[3 * so + 2 * C + p] = WV;
    }
    const CellOut<R> o =
        (VAR == 0) ? update_with_dta<R, VAR>(S, me.H, me.U, me.V, zbp, area, fnc, WH, WU, WV)
                   : update_math<R, VAR>(S, me.H, me.U, me.V, zbp, area, fnc, WH, WU, WV, sd[0], sd[1], sd[2], sd[3]);
    return State<R>{o.H, o.U, o.V, o.Z};
}

// One step of tile t under either scheme.
template <class R, int VAR, bool RARE, int SCHEME>
__device__ __forceinline__ void step_tile(const FusedView<R> &f, const State<R> *__restrict__ Sin,
                                          State<R> *__restrict__ Sout, int t, int kt, int jt) {
    if (SCHEME == 0) {
        tile_cells<R, VAR, RARE>(f, Sin, Sout, t, kt, jt);
    } else {
        const int p = t * f.block + threadIdx.x;
        if (p < f.C) fused_cell<R, VAR, RARE>(f, Sin, Sout, p, kt, jt);
    }
}

template <class R, int VAR, bool RARE, int SCHEME>
__global__ void __launch_bounds__(SCHEME == 0 ? kTileMaxBlock : kFusedMaxBlock) k_fused_step(FusedView<R> f, int kt,
                                                                                           int jt, int remap) {
    step_tile<R, VAR, RARE, SCHEME>(f, f.Sin, f.Sout, tile_of(blockIdx.x, f.ntiles, remap), kt, jt);
}

// Scenarios of one mesh in one launch: workgroup w runs tile w % ntiles of
// scenario w / ntiles.  The scenarios share the mesh's geometry and tasks and
// differ in state and Manning field (FNC [S][C]); each reads its own gauges.
template <class R, int VAR, bool RARE>
__global__ void __launch_bounds__(kTileMaxBlock) k_fused_step_scen(FusedView<R> f, int kt, int jt, int remap) {
    const int scen = blockIdx.x / f.ntiles, wg = blockIdx.x - scen * f.ntiles;
    const size_t so = (size_t)scen * f.C;
    tile_cells<R, VAR, RARE, false, true>(f, f.Sin + so, f.Sout + so, tile_of(wg, f.ntiles, remap), kt, jt, scen);
}

// TILES scheme limited to 128 VGPRs: four waves per SIMD (LDS allows four
// 256-thread workgroups per CU), for a 207K-cell mesh one resident round.
template <class R, int VAR, bool RARE>
__global__ void __launch_bounds__(kTileMaxBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_fused_step_o4(FusedView<R> f, int kt, int jt, int remap) {
    tile_cells<R, VAR, RARE>(f, f.Sin, f.Sout, tile_of(blockIdx.x, f.ntiles, remap), kt, jt);
}

// (day, kt) of global step g counted from day 0, kt 1 (benchmark.cu:89-90).
__device__ __forceinline__ void time_of(int64_t g, int spd, int *day, int *kt) {
    const int per = spd - 1;
    *day = (int)(g / per);
    *kt = 1 + (int)(g % per);
}

// The next step's (day, kt) from this step's, as the step loop counts them
// (kt in 1..spd-1): what time_of(g + 1) gives, without its 64-bit division.
__device__ __forceinline__ void time_next(int spd, int *day, int *kt) {
    if (++*kt >= spd) {
        *kt = 1;
        ++*day;
    }
}

// (day, kt) of the step `k` steps after the clock's (one uniform 32-bit
// division per wave).
__device__ __forceinline__ void clock_at(StepClock c, int k, int spd, int *day, int *kt) {
    const int per = spd - 1;
    const int t = c.kt - 1 + k;
    *day = c.day + t / per;
    *kt = 1 + t % per;
}

// Graph-replayable variant: the step's time comes from a device clock, so
// the same captured graph advances the boundary series on every replay (the
// reference's graph replays one frozen (kt=1, day=0) step, benchmark.cu:235-240).
template <class R, int VAR, bool RARE, int SCHEME>
__global__ void __launch_bounds__(SCHEME == 0 ? kTileMaxBlock : kFusedMaxBlock) k_fused_step_tick(FusedView<R> f,
                                                                                                const StepClock *clock,
                                                                    int offset, int spd, int remap) {
    int day, kt;
    clock_at(*clock, offset, spd, &day, &kt);
    step_tile<R, VAR, RARE, SCHEME>(f, f.Sin, f.Sout, tile_of(blockIdx.x, f.ntiles, remap), kt, day);
}

template <class R, int VAR, bool RARE>
__global__ void __launch_bounds__(kTileMaxBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_fused_step_tick_o4(FusedView<R> f, const StepClock *clock, int offset, int spd, int remap) {
    int day, kt;
    clock_at(*clock, offset, spd, &day, &kt);
    tile_cells<R, VAR, RARE>(f, f.Sin, f.Sout, tile_of(blockIdx.x, f.ntiles, remap), kt, day);
}

// The reference layout's flux kernel with the step's time from the device
// clock, for the graph-captured two-kernel mode.
template <class R, int VAR>
__global__ void __launch_bounds__(kMaxBlock) k_calculate_flux_tick(ExtView<R> m, const StepClock *clock, int offset,
                                                                   int spd) {
    int day, kt;
    clock_at(*clock, offset, spd, &day, &kt);
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < 4 * m.CELL) ext_flux_slot<R, VAR>(m, idx, kt, day);
}

__global__ void k_tick_advance(int64_t *tick, int by) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *tick += by;
}

__global__ void k_clock_advance(StepClock *clock, int by, int spd) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        int day, kt;
        clock_at(*clock, by, spd, &day, &kt);
        *clock = StepClock{day, kt};
    }
}

__global__ void k_clock_set(StepClock *clock, int day, int kt) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *clock = StepClock{day, kt};
}

// ---------------------------------------------------------------------------
// Persistent multi-step kernel.

// Grid-wide barrier: every wave releases its stores at device scope (write
// back to the point of coherence across XCDs), one lane per workgroup
// arrives on a monotonic counter and waits for `target`, then every wave
// acquires.  A wait longer than spin_limit sleeps sets *error and leaves, so
// a broken launch cannot hang the device.
__device__ __forceinline__ void grid_barrier(unsigned *counter, unsigned target, int *error,
                                             unsigned spin_limit) {
    // every wave waits until its own stores are complete (written through to
    // this XCD's L2), so one device-scope release per workgroup covers them
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        while (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > spin_limit) {
                __hip_atomic_store(error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// TILES scheme: no grid-wide barrier.  Tile t starts step s once every tile
// sharing an edge with it has finished step s-1 (its flag reached s): that
// covers both the halo it reads (their step s-1 output) and the buffer it
// overwrites (their step s-1 input, which they no longer read).  States move
// with device-coherent accesses, so publishing a tile is: every wave waits
// for its stores, then one relaxed flag store.  Neighbouring tiles stay
// within one step of each other; distant ones may drift further.
// Every workgroup takes tiles blockIdx.x, blockIdx.x + gridDim.x, ... in
// step-major order; a wait only ever points at an earlier step, so with all
// workgroups resident (cooperative launch) every wait ends.  A wait longer
// than spin_limit sleeps sets *error and the kernel leaves.
template <class R>
__device__ __forceinline__ bool wait_tiles(const FusedView<R> &f, int t, int target, const PersistentArgs &pa) {
    const int a = f.tile_nb_off[t], b = f.tile_nb_off[t + 1];
    int ok = 1;
    // one lane per neighbouring tile polls its flag back to back (a poll
    // is one device-coherent load, whose latency paces the loop); every 64th
    // poll also checks the spin limit and whether another tile gave up
    for (int k = a + (int)threadIdx.x; k < b; k += blockDim.x) {
        const int u = f.tile_nb[k];
        unsigned spins = 0;
        while (__hip_atomic_load(f.tile_flag + (size_t)u * kFlagStride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if ((++spins & 63) == 0 &&
                (spins > pa.spin_limit || __hip_atomic_load(pa.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                __hip_atomic_store(pa.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
        }
    }
    return __syncthreads_and(ok) != 0;
}

// The same wait by the lanes of one wavefront only (lane = thread index
// mod 64): no workgroup barrier, so the other wavefronts of the tile go on.
// Wavefront-uniform result.
template <class R>
__device__ __forceinline__ bool wave_wait_tiles(const FusedView<R> &f, int t, int target, const PersistentArgs &pa,
                                                int lane) {
    const int a = f.tile_nb_off[t], b = f.tile_nb_off[t + 1];
    bool ok = true;
    for (int k = a + lane; k < b; k += 64) {
        const int u = f.tile_nb[k];
        unsigned spins = 0;
        while (__hip_atomic_load(f.tile_flag + (size_t)u * kFlagStride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if ((++spins & 63) == 0 &&
                (spins > pa.spin_limit || __hip_atomic_load(pa.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                __hip_atomic_store(pa.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = false;
                break;
            }
        }
        if (!ok) break;
    }
    return __builtin_amdgcn_ballot_w64(!ok) == 0;
}

template <class R, int VAR, bool RARE>
__device__ __forceinline__ void persistent_tiles(const FusedView<R> &f, const PersistentArgs &pa) {
    const int ntiles = f.ntiles;
    const int64_t g0 = (int64_t)pa.day0 * (pa.spd - 1) + (pa.kt0 - 1);
    int day, kt;
    time_of(g0, pa.spd, &day, &kt);
    for (int s = 0; s < pa.nsteps; s++, time_next(pa.spd, &day, &kt)) {
        const State<R> *a = (s & 1) ? f.Sout : f.Sin;
        State<R> *b = (s & 1) ? const_cast<State<R> *>(f.Sin) : f.Sout;
        for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
            // the wait sits between the tile's own loads and its neighbours'
            if (!tile_cells<R, VAR, RARE, true>(f, a, b, t, kt, day, 0,
                                                [&]() { return s == 0 || wait_tiles(f, t, s, pa); }))
                return;
            __builtin_amdgcn_s_waitcnt(0);   // this wave's stores are complete
            __syncthreads();
            if (threadIdx.x == 0) __hip_atomic_store(f.tile_flag + (size_t)t * kFlagStride, s + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <class R, int VAR, bool RARE>
__global__ void __launch_bounds__(kTileMaxBlock) k_persistent_tiles(FusedView<R> f, PersistentArgs pa) {
    persistent_tiles<R, VAR, RARE>(f, pa);
}

// the same limited to 128 VGPRs (four workgroups per CU, some spills)
template <class R, int VAR, bool RARE>
__global__ void __launch_bounds__(kTileMaxBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_persistent_tiles_o4(FusedView<R> f, PersistentArgs pa) {
    persistent_tiles<R, VAR, RARE>(f, pa);
}

// Neighbour of a task in the resident kernel: a cell of the tile comes from
// LDS (its state of this step until the tile's update), another tile's from
// the coherent buffer `a`, which holds it for every halo cell.
template <class R>
__device__ __forceinline__ void resident_neighbour(const FusedView<R> &f, const TileLds<R> &L,
                                                   const State<R> *__restrict__ a, TaskIn<R> &t, int p0, int own) {
    const bool single = t.w != TASK_NONE && !(t.w & TASK_PAIR);
    const int ni = (single && t.nb >= 0) ? t.nb : own;
    const int l = ni - p0;
    if (l >= 0 && l < f.tile) {
        t.on = L.st[l];
        t.zbn = L.zb[l];
    } else {
        t.on = ld_state<true>(a, ni, f.C);
        t.zbn = f.ZBC[ni];
    }
}

template <class R> __device__ __forceinline__ void hold(TaskIn<R> &t) {
    asm volatile("" : "+v"(t.w), "+v"(t.nb), "+v"(t.c), "+v"(t.s), "+v"(t.sa));
    t.sb = t.sa;
}

// Resident TILES kernel: one workgroup per tile for the whole launch (the
// grid is exactly the tile count, all resident), so nothing a tile owns
// leaves the CU between steps.  The tile's statics and the task entries of
// its first two positions stay in registers, its cell states in LDS.  A step
// runs the chunks that read only the tile's own cells first (pairs, and
// one-slot tasks whose neighbour is in the tile, from LDS), then waits for
// the neighbouring tiles' previous step (wait_tiles, as persistent_tiles),
// then runs the chunks that read other tiles' halo cells (from the coherent
// ping-pong buffers), then the sums and update.  The new state goes into LDS
// for this tile's next step and, for the halo cells (and every cell at the
// last step), to the step's output buffer.  Same arithmetic as tile_cells,
// so the same bits.
// BLK: the workgroup size when fixed at compile time (its LDS offsets are
// then constants), 0 = f.block.
template <class R, int VAR, bool RARE, int BLK>
__device__ __forceinline__ void resident_tiles(const FusedView<R> &f, const PersistentArgs &pa, int remap) {
    const int t = tile_of(blockIdx.x, f.ntiles, remap);
    const int B = BLK ? BLK : f.block, C = f.C;
    const int p0 = t * f.tile;
    const TileLds<R> L = tile_lds<R>(B);
    R zbp, area, fnc;
    uint32_t pad;
    R sd[4] = {0, 0, 0, 0};
    TaskIn<R> t0 = none_task<R>(), t1 = none_task<R>();
    const int k0 = t * f.task_stride, kn = k0 + 64 * f.task_n[t];
    // the first npub wavefronts hold every cell other tiles read: once they
    // have stored their cells' new states the step is published
    const int npub = f.tile_pub[t];
    __shared__ unsigned lds_pub;   // publishing wavefronts counted in so far
    if (threadIdx.x == 0) lds_pub = 0;
    {
        const int i = threadIdx.x, p = p0 + i;
        const bool live = i < f.tile && p < C;
        const int pc = live ? p : C - 1;
        const State<R> me = ld_state<true>(f.Sin, pc, C);
        zbp = f.ZBC[pc];
        const R z1 = (f.ZB1 ? f.ZB1 : f.ZBC)[pc];
        area = (VAR == 0) ? f.DTA[pc] : f.AREA[pc];
        fnc = f.FNC[pc];
        pad = f.cell_pad[pc];
        if (VAR == 1) {
#pragma unroll
            for (int j = 0; j < 4; j++) sd[j] = f.SIDE[j * C + pc];
        }
        const int kw = k0 + (i & ~63);
        if (kw < kn) t0 = task_load<R>(f, k0 + i);
        if (kw + B < kn) t1 = task_load<R>(f, k0 + i + B);
        if (live) {
            L.st[i] = me;
            L.zb[i] = zbp;
            if (f.ZB1) L.zb1[i] = z1;
        }
    }
    // entries past the first two positions are loaded every step; their
    // wavefronts wait for the neighbouring tiles as those with halo chunks
    const bool extra = kn > k0 + 2 * B;
    const bool any_wait = __syncthreads_or((t0.w != TASK_NONE && (t0.w & TASK_HALO)) ||
                                           (t1.w != TASK_NONE && (t1.w & TASK_HALO)) || extra) != 0;
    // wavefront 0 waits when it has such chunks itself, or when no
    // wavefront of the tile has any
    const bool wait_any = !any_wait || __builtin_amdgcn_ballot_w64((t0.w != TASK_NONE && (t0.w & TASK_HALO)) ||
                                                                   (t1.w != TASK_NONE && (t1.w & TASK_HALO)) ||
                                                                   extra) != 0;
    const int64_t g0 = (int64_t)pa.day0 * (pa.spd - 1) + (pa.kt0 - 1);
    int day, kt;
    time_of(g0, pa.spd, &day, &kt);
    for (int s = 0; s < pa.nsteps; s++, time_next(pa.spd, &day, &kt)) {
        // The thread index and the values held across steps pass through an
        // empty asm every step, so nothing derived from them (LDS offsets,
        // the slots' SIDE*COSF products, ...) is hoisted out of the loop into
        // registers held across the whole step.
        int i = threadIdx.x;
        asm volatile("" : "+v"(i));
        hold(t0);
        hold(t1);
        asm volatile("" : "+v"(zbp), "+v"(area), "+v"(fnc), "+v"(pad));
        // chunks reading only the tile's own cells run before the wait, those
        // reading other tiles' cells (TASK_HALO) after it
        const bool pre0 = t0.w != TASK_NONE && !(t0.w & TASK_HALO);
        const bool pre1 = t1.w != TASK_NONE && !(t1.w & TASK_HALO);
        const bool post0 = t0.w != TASK_NONE && (t0.w & TASK_HALO);
        const bool post1 = t1.w != TASK_NONE && (t1.w & TASK_HALO);
        const int p = p0 + i;
        const bool live = i < f.tile && p < C;
        const int pc = live ? p : C - 1;
        const State<R> *a = (s & 1) ? f.Sout : f.Sin;
        State<R> *b = (s & 1) ? const_cast<State<R> *>(f.Sin) : f.Sout;
        // Wavefronts with chunks reading other tiles (or entries past the
        // first two positions) wait, on their own, until the neighbouring
        // tiles have finished step s-1: their states of step s are then in
        // `a`, and they no longer read `b` (their step s-1 input), which
        // this tile overwrites after the barrier below.  They wait as the
        // step starts (the plan gives them those chunks only), while the
        // others run the tile's own chunks.  A tile without such chunks lets
        // wavefront 0 make the same wait for `b`, after its own chunks.
        const bool halo_wave = __builtin_amdgcn_ballot_w64(post0 || post1 || extra) != 0;
        const bool waits = (i >> 6) == 0 ? wait_any : halo_wave;
        if (!halo_wave) {
            if (pre0) {
                if (!(t0.w & TASK_PAIR)) resident_neighbour(f, L, a, t0, p0, pc);
                task_run<R, VAR, RARE>(f, L, t0, a, B, p0, kt, day);
            }
            if (pre1) {
                if (!(t1.w & TASK_PAIR)) resident_neighbour(f, L, a, t1, p0, pc);
                task_run<R, VAR, RARE>(f, L, t1, a, B, p0, kt, day);
            }
        }
        bool ok = true;
        if (s > 0 && waits) ok = wave_wait_tiles(f, t, pa.flag0 + s, pa, i & 63);
        if (ok) {
            if (post0) {
                resident_neighbour(f, L, a, t0, p0, pc);
                task_run<R, VAR, RARE>(f, L, t0, a, B, p0, kt, day);
            }
            if (post1) {
                resident_neighbour(f, L, a, t1, p0, pc);
                task_run<R, VAR, RARE>(f, L, t1, a, B, p0, kt, day);
            }
            for (int k = k0 + i + 2 * B; k < kn; k += B) {
                TaskIn<R> tk = task_load<R>(f, k);
                resident_neighbour(f, L, a, tk, p0, pc);
                if (tk.w != TASK_NONE) task_run<R, VAR, RARE>(f, L, tk, a, B, p0, kt, day);
            }
            if (halo_wave) {   // own chunks of a wavefront that also reads other tiles
                if (pre0) {
                    if (!(t0.w & TASK_PAIR)) resident_neighbour(f, L, a, t0, p0, pc);
                    task_run<R, VAR, RARE>(f, L, t0, a, B, p0, kt, day);
                }
                if (pre1) {
                    if (!(t1.w & TASK_PAIR)) resident_neighbour(f, L, a, t1, p0, pc);
                    task_run<R, VAR, RARE>(f, L, t1, a, B, p0, kt, day);
                }
            }
        }
        // every task of the step has run: `a` is no longer read
        if (!__syncthreads_and(ok)) return;
        if (live) {
            const State<R> nw = tile_finish<R, VAR, RARE, false>(f, L, B, i, p, L.st[i], zbp, area, fnc, pad, sd, 0);
            // other tiles read only the halo cells; every cell at the end
            if ((pad & kCellHalo) || s == pa.nsteps - 1) st_state<true>(b, p, C, nw);
            L.st[i] = nw;
        }
        // The wavefronts holding the halo cells publish the step as soon as
        // their stores are complete: each counts itself in, and the last one
        // sets the flag (the tile's other cells are read by no other tile).
        if ((i >> 6) < npub) {
            __builtin_amdgcn_s_waitcnt(0);   // this wave's stores are complete
            if ((i & 63) == 0 &&
                __hip_atomic_fetch_add(&lds_pub, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) + 1 ==
                    (unsigned)(s + 1) * (unsigned)npub)
                __hip_atomic_store(f.tile_flag + (size_t)t * kFlagStride, pa.flag0 + s + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();   // LDS states and terms of the step are final
        if (npub == 0 && threadIdx.x == 0)
            __hip_atomic_store(f.tile_flag + (size_t)t * kFlagStride, pa.flag0 + s + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <class R, int VAR, bool RARE, int BLK>
__global__ void __launch_bounds__(kTileMaxBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_resident_tiles(FusedView<R> f, PersistentArgs pa, int remap) {
    resident_tiles<R, VAR, RARE, BLK>(f, pa, remap);
}

template <class R, int VAR, bool RARE, int SCHEME>
__global__ void __launch_bounds__(SCHEME == 0 ? kTileMaxBlock : kFusedMaxBlock) k_persistent(FusedView<R> f,
                                                                                           PersistentArgs pa) {
    const int ntiles = (f.C + blockDim.x - 1) / blockDim.x;
    const int64_t g0 = (int64_t)pa.day0 * (pa.spd - 1) + (pa.kt0 - 1);
    const State<R> *a = f.Sin;
    State<R> *b = f.Sout;
    int day, kt;
    time_of(g0, pa.spd, &day, &kt);
    for (int s = 0; s < pa.nsteps; s++, time_next(pa.spd, &day, &kt)) {
        for (int t = blockIdx.x; t < ntiles; t += gridDim.x) step_tile<R, VAR, RARE, SCHEME>(f, a, b, t, kt, day);
        State<R> *tmp = const_cast<State<R> *>(a);
        a = b;
        b = tmp;
        if (s + 1 < pa.nsteps) {
            grid_barrier(pa.barrier, (unsigned)(s + 1) * gridDim.x, pa.error, pa.spin_limit);
            if (__hip_atomic_load(pa.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
        }
    }
}

// ---------------------------------------------------------------------------
// Batches of meshes.

template <class R, int VAR, bool RARE, int SCHEME>
__device__ __forceinline__ void batch_tile(const BatchView<R> &b, int t, int step) {
    const int m = __builtin_amdgcn_readfirstlane(b.tile_mesh[t]);
    const FusedView<R> &f = b.views[m];
    int day, kt;
    time_of(b.g0[m] + step, b.spd[m], &day, &kt);
    const bool odd = step & 1;
    step_tile<R, VAR, RARE, SCHEME>(f, odd ? f.Sout : f.Sin, odd ? const_cast<State<R> *>(f.Sin) : f.Sout,
                                    t - b.tile0[m], kt, day);
}

template <class R, int VAR, bool RARE, int SCHEME>
__global__ void __launch_bounds__(kFusedMaxBlock) k_batch_step(BatchView<R> b, int step, int remap) {
    batch_tile<R, VAR, RARE, SCHEME>(b, tile_of(blockIdx.x, b.ntiles, remap), step);
}

template <class R, int VAR, bool RARE, int SCHEME>
__global__ void __launch_bounds__(kFusedMaxBlock) k_batch_step_tick(BatchView<R> b, const int64_t *tick, int offset,
                                                                    int remap) {
    batch_tile<R, VAR, RARE, SCHEME>(b, tile_of(blockIdx.x, b.ntiles, remap), (int)(*tick) + offset);
}

template <class R, int VAR, bool RARE, int SCHEME>
__global__ void __launch_bounds__(kFusedMaxBlock) k_batch_persistent(BatchView<R> b, PersistentArgs pa) {
    for (int s = 0; s < pa.nsteps; s++) {
        for (int t = blockIdx.x; t < b.ntiles; t += gridDim.x) batch_tile<R, VAR, RARE, SCHEME>(b, t, s);
        if (s + 1 < pa.nsteps) {
            grid_barrier(pa.barrier, (unsigned)(s + 1) * gridDim.x, pa.error, pa.spin_limit);
            if (__hip_atomic_load(pa.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
        }
    }
}

// ---------------------------------------------------------------------------
// Layout conversion.

template <class R>
__global__ void k_gather(ExtView<R> m, const int32_t *orig, State<R> *st, int C) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= C) return;
    const int o = orig[p];
    st[p] = State<R>{m.H[o], m.U[o], m.V[o], m.Z[o]};
}

template <class R>
__global__ void k_scatter(ExtView<R> m, const int32_t *orig, const State<R> *st, int C) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= C) return;
    const int o = orig[p];
    const State<R> s = st[p];
    m.H[o] = s.H;
    m.U[o] = s.U;
    m.V[o] = s.V;
    m.Z[o] = s.Z;
    m.W[o] = speed<R>(s.U, s.V);
}

__global__ void k_refresh_meta(const int32_t *KLAS_ext, const int32_t *orig, uint8_t *meta, int C) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 4 * C) return;
    const int j = e / C, p = e - j * C;
    const int mt = meta[e];
    if ((mt & META_KIND_MASK) == K_PAD) return;
    const bool prim = (mt & META_PRIMARY_BIT) != 0;
    meta[e] = (uint8_t)(kind_of_klas(KLAS_ext[4 * orig[p] + j], prim) | (prim ? META_PRIMARY_BIT : 0));
}

// DTA = DT / AREA is rebuilt with the expression the host used at prepare
// (1.0 * DT / (1.0 * AREA) in double, rounded to R), so the fused steps keep
// reading what the update would form from the new AREA.
template <class R>
__global__ void k_refresh_statics(ExtView<R> m, const int32_t *orig, R *ZBC, R *AREA, R *FNC, R *DTA, R *ZB1, int C) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= C) return;
    const int o = orig[p];
    ZBC[p] = m.ZBC[o];
    const R area = m.AREA[o];
    AREA[p] = area;
    DTA[p] = (R)((1.0 * (double)m.S.DT) / (1.0 * (double)area));
    FNC[p] = m.FNC[o];
    if (ZB1) ZB1[p] = m.ZB1[o];
}

inline int div_up(int a, int b) { return (a + b - 1) / b; }

// Regrouping of each tile's pair tasks by the wet/dry branches their two
// sides take in the current state (both wet, one side, neither; stable in
// entry order), as the host planner orders them: the pair tasks of tile t
// are permuted among the entries that hold pair tasks, so every chunk keeps
// its place, kind and size.  A pair task reads only its tile's LDS and
// writes its two slots' own terms, so which entry runs it changes no result.
// One workgroup of 1024 threads per tile; a tile has at most 2 pairs per
// cell (two of a cell's four slots each), so at most kMaxPairs.
constexpr int kMaxPairs = 2 * kTileMaxBlock;

template <class R>
__device__ __forceinline__ bool regroup_side_wet(const FusedView<R> &f, int a, int b) {
    const State<R> sa = f.Sin[a], sb = f.Sin[b];
    const R zba = f.ZBC[a], zbb = f.ZBC[b];
    const R zb1a = f.ZB1 ? f.ZB1[a] : (R)(zba + f.S.HM1);
    return ladder_branch<R>(f.S, sa.H, fmax(sa.Z, zb1a), zba, fmax(sb.H, f.S.HM1), zbb, fmax(zbb, sb.Z)) == 7;
}

template <class R> __device__ void regroup_tile(const FusedView<R> &f, int t) {
    __shared__ int32_t slot_of[kMaxPairs];   // entry of the r-th pair task in entry order
    __shared__ int32_t src_of[kMaxPairs];    // entry of the r-th pair task in grouped order
    __shared__ uint8_t cls_of[kMaxPairs];
    __shared__ int16_t crank_of[kMaxPairs];  // rank among the pairs of its class, in entry order
    __shared__ int32_t wcnt[16][4];          // per wavefront: pairs of class 0, 1, 2; all pairs
    __shared__ int32_t run[4];               // the same over the rounds so far
    const size_t k0 = (size_t)t * f.task_stride;
    const int n = 64 * f.task_n[t], p0 = t * f.tile;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (threadIdx.x < 4) run[threadIdx.x] = 0;
    __syncthreads();
    for (int base = 0; base < n; base += blockDim.x) {
        const int e = base + (int)threadIdx.x;
        const uint32_t w = e < n ? f.task_w[k0 + e] : TASK_NONE;
        const bool pair = w != TASK_NONE && (w & TASK_PAIR);
        int c = 3;
        if (pair) {
            const int a = p0 + (int)task_la(w), b = p0 + (int)task_lb(w);
            c = 2 - (int)regroup_side_wet(f, a, b) - (int)regroup_side_wet(f, b, a);
        }
        const uint64_t m0 = __builtin_amdgcn_ballot_w64(pair && c == 0);
        const uint64_t m1 = __builtin_amdgcn_ballot_w64(pair && c == 1);
        const uint64_t m2 = __builtin_amdgcn_ballot_w64(pair && c == 2);
        const uint64_t mp = __builtin_amdgcn_ballot_w64(pair);
        if (lane == 0) {
            wcnt[wv][0] = __popcll(m0);
            wcnt[wv][1] = __popcll(m1);
            wcnt[wv][2] = __popcll(m2);
            wcnt[wv][3] = __popcll(mp);
        }
        __syncthreads();
        if (pair) {
            const uint64_t below = (1ull << lane) - 1;
            const uint64_t mc = c == 0 ? m0 : c == 1 ? m1 : m2;
            int rc = run[c] + __popcll(mc & below), rp = run[3] + __popcll(mp & below);
            for (int v = 0; v < wv; v++) {
                rc += wcnt[v][c];
                rp += wcnt[v][3];
            }
            if (rp < kMaxPairs) {
                slot_of[rp] = e;
                cls_of[rp] = (uint8_t)c;
                crank_of[rp] = (int16_t)rc;
            }
        }
        __syncthreads();
        if (threadIdx.x < 4) {
            int sum = 0;
            for (int v = 0; v < nw; v++) sum += wcnt[v][threadIdx.x];
            run[threadIdx.x] += sum;
        }
        __syncthreads();
    }
    const int np = run[3];
    if (np > kMaxPairs) return;   // not a plan of this library: left as it is
    for (int r = threadIdx.x; r < np; r += blockDim.x) {
        const int c = cls_of[r];
        src_of[(c == 0 ? 0 : c == 1 ? run[0] : run[0] + run[1]) + crank_of[r]] = slot_of[r];
    }
    __syncthreads();
    // every thread holds at most two moved entries (np <= 2 * blockDim.x):
    // all are read before any is written
    uint32_t *tw = const_cast<uint32_t *>(f.task_w);
    int32_t *tnb = const_cast<int32_t *>(f.task_nb);
    R *tcs = const_cast<R *>(f.task_cs), *tsd = const_cast<R *>(f.task_sd);
    uint32_t w0 = 0, w1 = 0;
    int32_t nb0 = 0, nb1 = 0;
    R c0 = 0, s0 = 0, d0 = 0, c1 = 0, s1 = 0, d1 = 0;
    const int r0 = threadIdx.x, r1 = threadIdx.x + blockDim.x;
    if (r0 < np) {
        const size_t q = k0 + src_of[r0];
        w0 = tw[q]; nb0 = tnb[q]; c0 = tcs[2 * q]; s0 = tcs[2 * q + 1]; d0 = tsd[q];
    }
    if (r1 < np) {
        const size_t q = k0 + src_of[r1];
        w1 = tw[q]; nb1 = tnb[q]; c1 = tcs[2 * q]; s1 = tcs[2 * q + 1]; d1 = tsd[q];
    }
    __syncthreads();
    if (r0 < np) {
        const size_t q = k0 + slot_of[r0];
        tw[q] = w0; tnb[q] = nb0; tcs[2 * q] = c0; tcs[2 * q + 1] = s0; tsd[q] = d0;
    }
    if (r1 < np) {
        const size_t q = k0 + slot_of[r1];
        tw[q] = w1; tnb[q] = nb1; tcs[2 * q] = c1; tcs[2 * q + 1] = s1; tsd[q] = d1;
    }
}

template <class R> __global__ void __launch_bounds__(1024) k_regroup(FusedView<R> f) { regroup_tile(f, blockIdx.x); }

// The same for every tile of a batch (each member's view holds its current
// state in Sin).
template <class R> __global__ void __launch_bounds__(1024) k_batch_regroup(BatchView<R> b) {
    const int m = b.tile_mesh[blockIdx.x];
    const FusedView<R> f = b.views[m];
    regroup_tile(f, (int)blockIdx.x - b.tile0[m]);
}

}  // namespace

template <class R> hipError_t launch_regroup(const FusedView<R> &f, hipStream_t s) {
    if (f.scheme != 0 || f.ntiles == 0) return hipSuccess;   // TILES scheme only
    if (f.tile > kTileMaxBlock) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_regroup<R>), dim3(f.ntiles), dim3(1024), 0, s, f);
    return hipGetLastError();
}
template hipError_t launch_regroup<double>(const FusedView<double> &, hipStream_t);
template hipError_t launch_regroup<float>(const FusedView<float> &, hipStream_t);

template <class R> hipError_t launch_batch_regroup(const BatchView<R> &b, hipStream_t s) {
    if (b.scheme != 0 || b.ntiles == 0) return hipSuccess;   // TILES scheme only
    hipLaunchKernelGGL((k_batch_regroup<R>), dim3(b.ntiles), dim3(1024), 0, s, b);
    return hipGetLastError();
}
template hipError_t launch_batch_regroup<double>(const BatchView<double> &, hipStream_t);
template hipError_t launch_batch_regroup<float>(const BatchView<float> &, hipStream_t);

template <class R>
hipError_t launch_refresh_statics(const ExtView<R> &m, const int32_t *orig, R *ZBC, R *AREA, R *FNC, R *DTA, R *ZB1,
                                  int C, hipStream_t s) {
    if (C == 0) return hipSuccess;
    hipLaunchKernelGGL((k_refresh_statics<R>), dim3(div_up(C, 256)), dim3(256), 0, s, m, orig, ZBC, AREA, FNC, DTA, ZB1,
                       C);
    return hipGetLastError();
}
template hipError_t launch_refresh_statics<double>(const ExtView<double> &, const int32_t *, double *, double *,
                                                   double *, double *, double *, int, hipStream_t);
template hipError_t launch_refresh_statics<float>(const ExtView<float> &, const int32_t *, float *, float *, float *,
                                                  float *, float *, int, hipStream_t);

// ---------------------------------------------------------------------------
// Launch helpers.

template <class R, int VAR>
hipError_t launch_calculate_flux(const ExtView<R> &m, int kt, int jt, int block, hipStream_t s) {
    const int n = 4 * m.CELL;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL((k_calculate_flux<R, VAR>), dim3(div_up(n, block)), dim3(block), 0, s, m, kt, jt);
    return hipGetLastError();
}

template <class R, int VAR>
hipError_t launch_calculate_flux_tick(const ExtView<R> &m, const StepClock *clock, int offset, int spd, int block,
                                      hipStream_t s) {
    const int n = 4 * m.CELL;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL((k_calculate_flux_tick<R, VAR>), dim3(div_up(n, block)), dim3(block), 0, s, m, clock, offset,
                       spd);
    return hipGetLastError();
}

template <class R, int VAR>
hipError_t launch_update_cell(const ExtView<R> &m, int block, hipStream_t s) {
    if (m.CELL == 0) return hipSuccess;
    hipLaunchKernelGGL((k_update_cell<R, VAR>), dim3(div_up(m.CELL, block)), dim3(block), 0, s, m);
    return hipGetLastError();
}

template <class R> size_t step_lds(int scheme, int block) {
    return scheme == 0 ? tile_lds_bytes(block, (int)sizeof(R)) : 0;
}

// A launch with more than 64 KiB of dynamic LDS (TILES workgroups of more
// than 448 threads in fp64) must be allowed per kernel first; done once per
// kernel.
hipError_t allow_lds(const void *k, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    static std::mutex mu;
    static std::vector<std::pair<const void *, size_t>> done;
    std::lock_guard<std::mutex> g(mu);
    for (const auto &d : done)
        if (d.first == k && d.second >= bytes) return hipSuccess;
    const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess) done.push_back({k, bytes});
    return e;
}

// Largest workgroup of a scheme's single-mesh kernels (their launch bounds).
inline int max_block(int scheme) { return scheme == 0 ? kTileMaxBlock : kFusedMaxBlock; }

// Kernel pointer of a (rare, scheme) instantiation.
#define HYDRO_PICK(KERNEL, R, VAR, rare, scheme)                                                  \
    ((rare) ? ((scheme) == 0 ? (const void *)KERNEL<R, VAR, true, 0> : (const void *)KERNEL<R, VAR, true, 1>) \
            : ((scheme) == 0 ? (const void *)KERNEL<R, VAR, false, 0> : (const void *)KERNEL<R, VAR, false, 1>))

template <class R, int VAR>
hipError_t launch_fused_step(const FusedView<R> &f, int kt, int jt, int block, int remap, hipStream_t s) {
    if (f.C == 0) return hipSuccess;
    if (block > max_block(f.scheme) || block != f.block) return hipErrorInvalidValue;
    FusedView<R> a = f;
    void *args[] = {&a, &kt, &jt, &remap};
    const void *k = (f.occ4 && f.scheme == 0)
                        ? (f.rare ? (const void *)k_fused_step_o4<R, VAR, true> : (const void *)k_fused_step_o4<R, VAR, false>)
                        : HYDRO_PICK(k_fused_step, R, VAR, f.rare, f.scheme);
    const hipError_t e = allow_lds(k, step_lds<R>(f.scheme, block));
    if (e != hipSuccess) return e;
    return hipLaunchKernel(k, dim3(f.ntiles), dim3(block), args, step_lds<R>(f.scheme, block), s);
}

template <class R, int VAR>
hipError_t launch_fused_step_scen(const FusedView<R> &f, int nscen, int kt, int jt, int block, int remap,
                                  hipStream_t s) {
    if (f.C == 0 || nscen <= 0) return hipSuccess;
    if (block > kTileMaxBlock || block != f.block || f.scheme != 0 || f.rare) return hipErrorInvalidValue;
    FusedView<R> a = f;
    void *args[] = {&a, &kt, &jt, &remap};
    const hipError_t e = allow_lds((const void *)k_fused_step_scen<R, VAR, false>, step_lds<R>(f.scheme, block));
    if (e != hipSuccess) return e;
    return hipLaunchKernel((const void *)k_fused_step_scen<R, VAR, false>, dim3(f.ntiles * nscen), dim3(block), args,
                           step_lds<R>(f.scheme, block), s);
}

template <class R, int VAR>
hipError_t launch_fused_step_tick(const FusedView<R> &f, const StepClock *clock, int offset, int spd, int block,
                                  int remap, hipStream_t s) {
    if (f.C == 0) return hipSuccess;
    if (block > max_block(f.scheme) || block != f.block) return hipErrorInvalidValue;
    FusedView<R> a = f;
    void *args[] = {&a, &clock, &offset, &spd, &remap};
    const void *k = (f.occ4 && f.scheme == 0)
                        ? (f.rare ? (const void *)k_fused_step_tick_o4<R, VAR, true>
                                  : (const void *)k_fused_step_tick_o4<R, VAR, false>)
                        : HYDRO_PICK(k_fused_step_tick, R, VAR, f.rare, f.scheme);
    const hipError_t e = allow_lds(k, step_lds<R>(f.scheme, block));
    if (e != hipSuccess) return e;
    return hipLaunchKernel(k, dim3(f.ntiles), dim3(block), args, step_lds<R>(f.scheme, block), s);
}

__global__ void k_fill_i32(int32_t *p, int32_t v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_i32(int32_t *p, int32_t v, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3(div_up(n, 256)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_tick_advance(int64_t *tick, int by, hipStream_t s) {
    hipLaunchKernelGGL(k_tick_advance, dim3(1), dim3(64), 0, s, tick, by);
    return hipGetLastError();
}

hipError_t launch_clock_advance(StepClock *clock, int by, int spd, hipStream_t s) {
    hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(64), 0, s, clock, by, spd);
    return hipGetLastError();
}

hipError_t launch_clock_set(StepClock *clock, int day, int kt, hipStream_t s) {
    hipLaunchKernelGGL(k_clock_set, dim3(1), dim3(64), 0, s, clock, day, kt);
    return hipGetLastError();
}

template <class R, int VAR> const void *persistent_kernel(bool rare, int scheme, bool occ4) {
    if (scheme == 0 && occ4)
        return rare ? (const void *)k_persistent_tiles_o4<R, VAR, true> : (const void *)k_persistent_tiles_o4<R, VAR, false>;
    if (scheme == 0)
        return rare ? (const void *)k_persistent_tiles<R, VAR, true> : (const void *)k_persistent_tiles<R, VAR, false>;
    return HYDRO_PICK(k_persistent, R, VAR, rare, scheme);
}

template <class R, int VAR>
hipError_t persistent_max_grid(int block, bool rare, int scheme, bool occ4, int *grid) {
    int dev = 0, per_cu = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    e = allow_lds(persistent_kernel<R, VAR>(rare, scheme, occ4), step_lds<R>(scheme, block));
    if (e != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persistent_kernel<R, VAR>(rare, scheme, occ4), block,
                                                     step_lds<R>(scheme, block));
    if (e != hipSuccess) return e;
    *grid = per_cu * cus;
    return hipSuccess;
}

template <class R, int VAR>
hipError_t launch_persistent(const FusedView<R> &f, const PersistentArgs &pa, int block, int grid, hipStream_t s) {
    if (block > max_block(f.scheme) || block != f.block) return hipErrorInvalidValue;
    FusedView<R> a = f;
    PersistentArgs p = pa;
    void *args[] = {&a, &p};
    const hipError_t e = allow_lds(persistent_kernel<R, VAR>(f.rare != 0, f.scheme, f.occ4 != 0),
                                   step_lds<R>(f.scheme, block));
    if (e != hipSuccess) return e;
    return hipLaunchCooperativeKernel(persistent_kernel<R, VAR>(f.rare != 0, f.scheme, f.occ4 != 0), dim3(grid),
                                      dim3(block), args,
                                      step_lds<R>(f.scheme, block), s);
}

// the 1024-thread build for 1024-thread workgroups, else the general one
template <class R, int VAR> const void *resident_kernel(bool rare, int block) {
    if (block == 1024)
        return rare ? (const void *)k_resident_tiles<R, VAR, true, 1024> : (const void *)k_resident_tiles<R, VAR, false, 1024>;
    return rare ? (const void *)k_resident_tiles<R, VAR, true, 0> : (const void *)k_resident_tiles<R, VAR, false, 0>;
}

template <class R, int VAR>
hipError_t resident_capacity(int block, bool rare, int *tiles) {
    int dev = 0, per_cu = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    e = allow_lds(resident_kernel<R, VAR>(rare, block), step_lds<R>(0, block));
    if (e != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, resident_kernel<R, VAR>(rare, block), block,
                                                     step_lds<R>(0, block));
    if (e != hipSuccess) return e;
    *tiles = per_cu * cus;
    return hipSuccess;
}

// plain: an ordinary launch; the caller has checked that every tile fits on
// the GPU at once and that nothing else holds CUs meanwhile.
template <class R, int VAR>
hipError_t launch_resident(const FusedView<R> &f, const PersistentArgs &pa, int remap, bool plain, hipStream_t s) {
    if (f.block > kTileMaxBlock || f.scheme != 0 || f.ntiles < 1) return hipErrorInvalidValue;
    FusedView<R> a = f;
    PersistentArgs p = pa;
    void *args[] = {&a, &p, &remap};
    const hipError_t e = allow_lds(resident_kernel<R, VAR>(f.rare != 0, f.block), step_lds<R>(0, f.block));
    if (e != hipSuccess) return e;
    if (plain)
        return hipLaunchKernel(resident_kernel<R, VAR>(f.rare != 0, f.block), dim3(f.ntiles), dim3(f.block), args,
                               step_lds<R>(0, f.block), s);
    return hipLaunchCooperativeKernel(resident_kernel<R, VAR>(f.rare != 0, f.block), dim3(f.ntiles), dim3(f.block), args,
                                      step_lds<R>(0, f.block), s);
}

template <class R>
hipError_t launch_gather(const ExtView<R> &m, const int32_t *orig, State<R> *st, int C, hipStream_t s) {
    if (C == 0) return hipSuccess;
    hipLaunchKernelGGL((k_gather<R>), dim3(div_up(C, 256)), dim3(256), 0, s, m, orig, st, C);
    return hipGetLastError();
}

template <class R>
hipError_t launch_scatter(const ExtView<R> &m, const int32_t *orig, const State<R> *st, int C, hipStream_t s) {
    if (C == 0) return hipSuccess;
    hipLaunchKernelGGL((k_scatter<R>), dim3(div_up(C, 256)), dim3(256), 0, s, m, orig, st, C);
    return hipGetLastError();
}

hipError_t launch_refresh_meta(const int32_t *KLAS_ext, const int32_t *orig, uint8_t *meta, int C,
                               hipStream_t s) {
    if (C == 0) return hipSuccess;
    hipLaunchKernelGGL(k_refresh_meta, dim3(div_up(4 * C, 256)), dim3(256), 0, s, KLAS_ext, orig, meta, C);
    return hipGetLastError();
}

template <class R, int VAR>
hipError_t launch_batch_step(const BatchView<R> &b, int step, int block, int remap, hipStream_t s) {
    if (b.ntiles == 0) return hipSuccess;
    if (block > kFusedMaxBlock) return hipErrorInvalidValue;
    BatchView<R> bb = b;
    void *args[] = {&bb, &step, &remap};
    return hipLaunchKernel(HYDRO_PICK(k_batch_step, R, VAR, b.rare, b.scheme), dim3(b.ntiles), dim3(block), args,
                           step_lds<R>(b.scheme, block), s);
}

template <class R, int VAR>
hipError_t launch_batch_step_tick(const BatchView<R> &b, const int64_t *tick, int offset, int block, int remap,
                                  hipStream_t s) {
    if (b.ntiles == 0) return hipSuccess;
    if (block > kFusedMaxBlock) return hipErrorInvalidValue;
    BatchView<R> bb = b;
    void *args[] = {&bb, &tick, &offset, &remap};
    return hipLaunchKernel(HYDRO_PICK(k_batch_step_tick, R, VAR, b.rare, b.scheme), dim3(b.ntiles), dim3(block),
                           args, step_lds<R>(b.scheme, block), s);
}

template <class R, int VAR>
hipError_t batch_persistent_max_grid(int block, bool rare, int scheme, int *grid) {
    int dev = 0, per_cu = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, HYDRO_PICK(k_batch_persistent, R, VAR, rare, scheme),
                                                     block, step_lds<R>(scheme, block));
    if (e != hipSuccess) return e;
    *grid = per_cu * cus;
    return hipSuccess;
}

template <class R, int VAR>
hipError_t launch_batch_persistent(const BatchView<R> &b, const PersistentArgs &pa, int block, int grid,
                                   hipStream_t s) {
    if (block > kFusedMaxBlock) return hipErrorInvalidValue;
    BatchView<R> bb = b;
    PersistentArgs p = pa;
    void *args[] = {&bb, &p};
    return hipLaunchCooperativeKernel(HYDRO_PICK(k_batch_persistent, R, VAR, b.rare != 0, b.scheme), dim3(grid),
                                      dim3(block), args, step_lds<R>(b.scheme, block), s);
}

#define HYDRO_INSTANTIATE(R, VAR)                                                                        \
    template hipError_t launch_calculate_flux<R, VAR>(const ExtView<R> &, int, int, int, hipStream_t);   \
    template hipError_t launch_fused_step_scen<R, VAR>(const FusedView<R> &, int, int, int, int, int, hipStream_t);  \
    template hipError_t launch_calculate_flux_tick<R, VAR>(const ExtView<R> &, const StepClock *, int, int, int,  \
                                                           hipStream_t);                                         \
    template hipError_t launch_update_cell<R, VAR>(const ExtView<R> &, int, hipStream_t);               \
    template hipError_t launch_fused_step<R, VAR>(const FusedView<R> &, int, int, int, int, hipStream_t); \
    template hipError_t launch_fused_step_tick<R, VAR>(const FusedView<R> &, const StepClock *, int, int, \
                                                       int, int, hipStream_t);                          \
    template hipError_t launch_persistent<R, VAR>(const FusedView<R> &, const PersistentArgs &, int, int, \
                                                  hipStream_t);                                         \
    template hipError_t persistent_max_grid<R, VAR>(int, bool, int, bool, int *);                                     \
    template hipError_t resident_capacity<R, VAR>(int, bool, int *);                                   \
    template hipError_t launch_resident<R, VAR>(const FusedView<R> &, const PersistentArgs &, int, bool, hipStream_t); \
    template hipError_t launch_batch_step<R, VAR>(const BatchView<R> &, int, int, int, hipStream_t);   \
    template hipError_t launch_batch_step_tick<R, VAR>(const BatchView<R> &, const int64_t *, int, int, int, \
                                                       hipStream_t);                                    \
    template hipError_t launch_batch_persistent<R, VAR>(const BatchView<R> &, const PersistentArgs &, int, int, \
                                                        hipStream_t);                                   \
    template hipError_t batch_persistent_max_grid<R, VAR>(int, bool, int, int *);

HYDRO_INSTANTIATE(double, 0)
HYDRO_INSTANTIATE(double, 1)
HYDRO_INSTANTIATE(float, 0)
HYDRO_INSTANTIATE(float, 1)

template hipError_t launch_gather<double>(const ExtView<double> &, const int32_t *, State<double> *, int,
                                          hipStream_t);
template hipError_t launch_gather<float>(const ExtView<float> &, const int32_t *, State<float> *, int,
                                         hipStream_t);
template hipError_t launch_scatter<double>(const ExtView<double> &, const int32_t *, const State<double> *, int,
                                           hipStream_t);
template hipError_t launch_scatter<float>(const ExtView<float> &, const int32_t *, const State<float> *, int,
                                          hipStream_t);

}  // namespace hydro
