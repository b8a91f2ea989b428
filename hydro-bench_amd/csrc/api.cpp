// api.cpp -- the C ABI declared in include/hydro.h.
#include "hydro.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "diag.hpp"
#include "kernels.hpp"
#include "prep.hpp"

using namespace hydro;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(...)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (__VA_ARGS__);                                                          \
        if (e_ != hipSuccess)                                                                   \
            return fail(HYDRO_ERR_HIP, "%s failed: %s (%s:%d)", #__VA_ARGS__,                   \
                        hipGetErrorString(e_), __FILE__, __LINE__);                             \
    } while (0)

// Device buffer owned by a mesh handle.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t n) {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = n;
        if (n == 0) return HYDRO_OK;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) return fail(HYDRO_ERR_NOMEM, "hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
        return HYDRO_OK;
    }
    template <class T> T *as() const { return (T *)p; }
};

struct GraphKey {
    int mode, parity, chunk;
    bool operator<(const GraphKey &o) const {
        if (mode != o.mode) return mode < o.mode;
        if (parity != o.parity) return parity < o.parity;
        return chunk < o.chunk;
    }
};

struct GraphExec {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    ~GraphExec() {
        if (x) (void)hipGraphExecDestroy(x);
        if (g) (void)hipGraphDestroy(g);
    }
};

bool is_device_ptr(const void *p) {
    if (!p) return false;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged || a.type == hipMemoryTypeUnified;
}

}  // namespace

struct hydro_mesh {
    hydro_mesh_desc d{};
    int spd = 0;          // steps_per_day = (int)(MDT / DT)
    int day = 0, kt = 1;  // time cursor: the next step runs at (kt, day)
    int CELL = 0;
    int rsize = 8;
    hydro_tune tune{};
    int block_two = 256;

    // kinds present (from KLAS at create time; KLAS codes are fixed for the
    // mesh's life except KLAS 7 -> 0, which only removes a kind)
    bool has_kind[16] = {};
    bool rare = false;   // a KLAS 3/5/6/7 or unknown code is present

    // reordered layout
    bool prepared = false;
    int cur = 0;  // state record buffer holding the current state
    bool ext_valid = true, int_valid = false;
    DevBuf st[2], ZBC, AREA, FNC, DTA, ZB1, nbr, meta, COSF, SINF, SIDE, SLCOS, SLSIN, orig, inv, twin;
    DevBuf twin_kind;   // [4][C]: kind of each slot's twin slot (K_PAD none), read by the fused adjoint step
    DevBuf task_w, task_nb, task_cs, task_sd, task_n, cell_pad;   // TILES scheme task entries
    DevBuf tile_nb_off, tile_nb, tile_flag;   // persistent TILES kernel: neighbour tiles (CSR), step flags
    DevBuf tile_pub;                          // resident kernel: wavefronts holding a tile's halo cells
    int max_tile_nb = 0;
    int tile = 256;                 // cells per tile (resolved hydro_tune.tile_cells)
    bool persist_pending = false;   // a persistent launch whose wait flag has not been read back
    bool last_resident = false;     // the last PERSISTENT launch ran the resident TILES kernel
    int resident_cap = -1;          // tiles the resident kernel holds at once (-1: not yet asked)
    int flag_base = 0;              // value every tile step flag holds (between persistent TILES launches)
    int since_group = 0;            // steps since the task plan was last grouped by the state
    bool defer_check = false;       // later launches of one request: leave the persistent check pending
    int task_stride = 0;
    int n_pairs = 0, n_singles = 0;    // tasks of the plan (statistics)
    DevBuf wbar, adj_flag;   // adjoint scratch: flux-sum adjoints [3][C], unsupported-kind flag
    DevBuf adj_ub;           // misfit sweeps: reversed cell updates, [2][S][kUbDoubles][C] doubles
    // misfit sweeps: the side stream the recomputation runs on, and the
    // events ordering it with the adjoint steps (created on first use)
    hipStream_t side_stream = nullptr;
    hipEvent_t sweep_start = nullptr, rec_done[2] = {}, adj_done[2] = {};
    DevBuf scen_st[2], scen_fnc;   // scenario misfit loops: states [S][C] (ping-pong), FNC [S][C]
    DevBuf gmap, gnext, gcells;   // misfit loops: gauge map [C], chain [n], internal ids [n]
    DevBuf tick, barrier, perr, ext_dev;   // tick: StepClock of the GRAPH mode
    DevBuf diag_part;   // hydro_diagnostics: one partial record per 1024 cells
    // Real-typed NAC / KLAS of the caller (index_real): int32 copies owned
    // here serve every kernel (d.NAC / d.KLAS point at them); klas_real is
    // the caller's Real KLAS, which KLAS 7 rewrites also update
    DevBuf nac_own, klas_own;
    void *klas_real = nullptr;
    bool need_zb1 = false, need_slcs = false;
    std::vector<double> xc, yc;   // cell coordinates (hydro_set_cell_coords), caller order
    hipStream_t cap_stream = nullptr;
    std::map<GraphKey, std::unique_ptr<GraphExec>> graphs;

    // profiling
    bool profiling = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0;
    int last_launches = 0, last_steps = 0;

    ~hydro_mesh() {
        graphs.clear();
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        if (side_stream) (void)hipStreamDestroy(side_stream);
        for (hipEvent_t e : {sweep_start, rec_done[0], rec_done[1], adj_done[0], adj_done[1]})
            if (e) (void)hipEventDestroy(e);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }
};

namespace {

template <class R> ExtView<R> ext_view(const hydro_mesh *m) {
    const hydro_mesh_desc &d = m->d;
    ExtView<R> v;
    v.CELL = d.CELL;
    v.NDAYS = d.NDAYS;
    v.NHQ = d.NHQ;
    v.S.DT = (R)d.DT;
    v.S.HM1 = (R)d.HM1;
    v.S.HM2 = (R)d.HM2;
    v.S.JL = (R)d.JL;
    v.H = (R *)d.H; v.U = (R *)d.U; v.V = (R *)d.V; v.Z = (R *)d.Z; v.W = (R *)d.W;
    v.ZBC = (const R *)d.ZBC; v.ZB1 = (const R *)d.ZB1; v.AREA = (const R *)d.AREA;
    v.FNC = (const R *)d.FNC; v.BF = (const R *)d.BoundaryFeature; v.NHQ1 = d.NHQ1;
    v.ZW = (const R *)d.ZW; v.QW = (const R *)d.QW; v.QT = (const R *)d.QT; v.DQT = (const R *)d.DQT;
    v.ZT = (const R *)d.ZT; v.DZT = (const R *)d.DZT;
    v.NAC = (const int32_t *)d.NAC; v.KLAS = (int32_t *)d.KLAS;
    v.KLASr = (R *)m->klas_real;
    v.SIDE = (const R *)d.SIDE; v.COSF = (const R *)d.COSF; v.SINF = (const R *)d.SINF;
    v.SLCOS = (const R *)d.SLCOS; v.SLSIN = (const R *)d.SLSIN;
    v.F0 = (R *)d.FLUX0; v.F1 = (R *)d.FLUX1; v.F2 = (R *)d.FLUX2; v.F3 = (R *)d.FLUX3;
    return v;
}

template <class R> FusedView<R> fused_view(const hydro_mesh *m, int in) {
    FusedView<R> f;
    f.C = m->CELL;
    f.S.DT = (R)m->d.DT;
    f.S.HM1 = (R)m->d.HM1;
    f.S.HM2 = (R)m->d.HM2;
    f.S.JL = (R)m->d.JL;
    f.ext = m->ext_dev.as<ExtView<R>>();
    f.Sin = m->st[in].as<State<R>>();
    f.Sout = m->st[in ^ 1].as<State<R>>();
    f.ZBC = m->ZBC.as<R>();
    f.AREA = m->AREA.as<R>();
    f.FNC = m->FNC.as<R>();
    f.DTA = m->DTA.as<R>();
    f.ZB1 = m->need_zb1 ? m->ZB1.as<R>() : nullptr;
    f.nbr = m->nbr.as<int32_t>();
    f.meta = m->meta.as<uint8_t>();
    f.COSF = m->COSF.as<R>();
    f.SINF = m->SINF.as<R>();
    f.SIDE = m->SIDE.as<R>();
    f.SLCOS = m->need_slcs ? m->SLCOS.as<R>() : nullptr;
    f.SLSIN = m->need_slcs ? m->SLSIN.as<R>() : nullptr;
    f.orig = m->orig.as<int32_t>();
    f.task_stride = m->task_stride;
    f.task_w = m->task_w.as<uint32_t>();
    f.task_nb = m->task_nb.as<int32_t>();
    f.task_cs = m->task_cs.as<R>();
    f.task_sd = m->task_sd.as<R>();
    f.task_n = m->task_n.as<int32_t>();
    f.cell_pad = m->cell_pad.as<uint8_t>();
    f.materialize = m->tune.materialize_flux;
    f.scheme = m->tune.scheme;
    f.block = m->tune.block;
    f.tile = m->tile;
    f.ntiles = (m->CELL + m->tile - 1) / m->tile;
    f.occ4 = m->tune.waves_per_simd == 4;
    f.rare = m->rare;
    f.rec_state = nullptr;
    f.rec_wsum = nullptr;
    f.tile_nb_off = m->tile_nb_off.as<int32_t>();
    f.tile_nb = m->tile_nb.as<int32_t>();
    f.tile_flag = m->tile_flag.as<int32_t>();
    f.tile_pub = m->tile_pub.as<int32_t>();
    f.gauge_map = nullptr;
    f.gauge_next = nullptr;
    f.gauge_n = 0;
    f.gauge_out = nullptr;
    return f;
}

template <class T> int d2h(std::vector<T> &dst, const void *src, size_t n) {
    dst.resize(n);
    if (n == 0) return HYDRO_OK;
    HIP_TRY(hipMemcpy(dst.data(), src, n * sizeof(T), hipMemcpyDeviceToHost));
    return HYDRO_OK;
}

template <class T> int h2d(DevBuf &b, const std::vector<T> &src) {
    int rc = b.alloc(src.size() * sizeof(T));
    if (rc) return rc;
    if (!src.empty()) HIP_TRY(hipMemcpy(b.p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return HYDRO_OK;
}

template <class R> int prepare_t(hydro_mesh *m) {
    const hydro_mesh_desc &d = m->d;
    const int C = d.CELL;
    const size_t S4 = 4 * (size_t)C;
    std::vector<int32_t> NAC, KLAS;
    std::vector<R> ZBC, ZB1, AREA, FNC, SIDE, COSF, SINF, SLCOS, SLSIN;
    int rc;
    if ((rc = d2h(NAC, (const int32_t *)d.NAC, S4)) || (rc = d2h(KLAS, (const int32_t *)d.KLAS, S4)) ||
        (rc = d2h(ZBC, d.ZBC, C)) ||
        (rc = d2h(ZB1, d.ZB1, C)) || (rc = d2h(AREA, d.AREA, C)) || (rc = d2h(FNC, d.FNC, C)) ||
        (rc = d2h(SIDE, d.SIDE, S4)) || (rc = d2h(COSF, d.COSF, S4)) || (rc = d2h(SINF, d.SINF, S4)) ||
        (rc = d2h(SLCOS, d.SLCOS, S4)) || (rc = d2h(SLSIN, d.SLSIN, S4)))
        return rc;

    std::vector<int32_t> perm;
    if (m->tune.reorder == HYDRO_REORDER_RCM) {
        perm = rcm_order(C, NAC.data());
    } else if (m->tune.reorder == HYDRO_REORDER_TILES) {
        perm = tile_order(C, NAC.data(), m->tile, rcm_order(C, NAC.data()));
    } else if (m->tune.reorder == HYDRO_REORDER_NONE) {
        perm.resize(C);
        for (int i = 0; i < C; i++) perm[i] = i;
    } else if (m->tune.reorder == HYDRO_REORDER_HILBERT) {
        if ((int)m->xc.size() != C)
            return fail(HYDRO_ERR_UNSUPPORTED, "the Hilbert order needs cell coordinates: call hydro_set_cell_coords");
        perm = hilbert_order(C, m->xc.data(), m->yc.data());
    } else {
        return fail(HYDRO_ERR_INVALID, "reorder %d", m->tune.reorder);
    }
    // the state at planning time (H, Z): cells and tasks are grouped by the
    // wet/dry branches they take in it
    std::vector<R> Hs, Zs;
    if ((rc = d2h(Hs, d.H, C)) || (rc = d2h(Zs, d.Z, C))) return rc;
    // TILES scheme: a tile's cells read by other tiles first, so the
    // wavefronts holding them finish their update, and publish it, first;
    // within each group the deep cells first, so the shallow and dry ones
    // share wavefronts (an update skips the friction terms when every cell
    // of its wavefront is shallow)
    if (m->tune.scheme == HYDRO_SCHEME_TILES && m->tune.reorder != HYDRO_REORDER_NONE) {
        std::vector<uint8_t> shallow(C);
        for (int o = 0; o < C; o++) shallow[o] = Hs[o] <= (R)d.HM2;
        edge_cells_first(C, NAC.data(), m->tile, perm, shallow.data());
    }
    std::vector<int32_t> inv(C);
    for (int p = 0; p < C; p++) inv[perm[p]] = p;

    // Are ZB1 and SLCOS/SLSIN exactly what the kernel would recompute?
    const R HM1 = (R)d.HM1;
    bool zb1_ok = true, slcs_ok = true;
    for (int i = 0; i < C && zb1_ok; i++) {
        const R z = ZBC[i] + HM1;
        zb1_ok = std::memcmp(&ZB1[i], &z, sizeof(R)) == 0;
    }
    for (size_t s = 0; s < S4 && slcs_ok; s++) {
        R a = SIDE[s] * COSF[s], b = SIDE[s] * SINF[s];
        slcs_ok = std::memcmp(&SLCOS[s], &a, sizeof(R)) == 0 && std::memcmp(&SLSIN[s], &b, sizeof(R)) == 0;
    }
    m->need_zb1 = !zb1_ok;
    m->need_slcs = !slcs_ok;

    // the same state in the new order (H, Z, ZB1), for the task grouping
    std::vector<R> h_n(C), z_n(C), zi_n(C);
    for (int p = 0; p < C; p++) {
        const int o = perm[p];
        h_n[p] = Hs[o];
        z_n[p] = Zs[o];
        zi_n[p] = std::max(Zs[o], ZB1[o]);
    }
    std::vector<R> zbc_n(C), area_n(C), fnc_n(C), zb1_n(m->need_zb1 ? C : 0), dta_n(C);
    std::vector<int32_t> nbr_n(S4), orig_n(C);
    std::vector<uint8_t> meta_n(S4);
    std::vector<R> cos_n(S4), sin_n(S4), side_n(S4), slc_n(m->need_slcs ? S4 : 0), sls_n(m->need_slcs ? S4 : 0);
    std::vector<uint8_t> twin_n(S4, 0xff);
    for (int p = 0; p < C; p++) {
        const int o = perm[p];
        orig_n[p] = o;
        zbc_n[p] = ZBC[o];
        area_n[p] = AREA[o];
        // DT / AREA exactly as the update forms it (1.0 * DT / (1.0 * AREA) in double)
        dta_n[p] = (R)((1.0 * (double)(R)d.DT) / (1.0 * (double)AREA[o]));
        fnc_n[p] = FNC[o];
        if (m->need_zb1) zb1_n[p] = ZB1[o];
        for (int j = 0; j < 4; j++) {
            const size_t so = 4 * (size_t)o + j, sn = (size_t)j * C + p;
            const int NC = NAC[so] - 1;
            const bool prim = (o < NC);   // decided on ORIGINAL ids (F:151)
            const R zero = 0;
            const bool pad = NAC[so] == 0 && KLAS[so] == 0 &&
                             std::memcmp(&SIDE[so], &zero, sizeof(R)) == 0 &&
                             std::memcmp(&COSF[so], &zero, sizeof(R)) == 0 &&
                             std::memcmp(&SINF[so], &zero, sizeof(R)) == 0 &&
                             std::memcmp(&SLCOS[so], &zero, sizeof(R)) == 0 &&
                             std::memcmp(&SLSIN[so], &zero, sizeof(R)) == 0;
            int kind = pad ? K_PAD : kind_of_klas(KLAS[so], prim);
            meta_n[sn] = (uint8_t)(kind | (prim ? META_PRIMARY_BIT : 0));
            nbr_n[sn] = (NC >= 0) ? inv[NC] : -1;
            if (NC >= 0)
                for (int j2 = 0; j2 < 4; j2++)
                    if (NAC[4 * (size_t)NC + j2] - 1 == o) {
                        twin_n[sn] = (uint8_t)j2;
                        break;
                    }
            cos_n[sn] = COSF[so];
            sin_n[sn] = SINF[so];
            side_n[sn] = SIDE[so];
            if (m->need_slcs) {
                slc_n[sn] = SLCOS[so];
                sls_n[sn] = SLSIN[so];
            }
        }
    }
    // TILES scheme tasks.  An interior edge whose two slots lie in one tile,
    // one primary and one mirror, with exactly opposite normals, is one pair
    // task (listed from its primary slot); every other non-padding slot is a
    // one-slot task.
    //
    // Layout: thread i of the workgroup runs entries i, i + block, i + 2 block,
    // ... of its tile, so each wavefront runs a column of 64-entry chunks.  A
    // wavefront pays for every code path any of its lanes takes, so chunks
    // are homogeneous: pair chunks, and one-slot chunks (interior slots, then
    // boundary slots).  The leftover pairs of a tile (its pair count modulo
    // 64) become two one-slot tasks each when they fit in the free entries
    // of the last one-slot chunk, which costs nothing there and saves a
    // mostly empty pair chunk.  Chunks are placed as described below, so
    // the wavefronts of a workgroup carry near-equal work and the chunks
    // that need other tiles' states come last.  Every tile gets the same
    // number of entries (its rounds times the block, the largest over
    // tiles), so a tile's entries start at t * stride.
    {
        const int B = m->tile;
        const int ntiles = (C + B - 1) / B;
        std::vector<std::vector<uint32_t>> lists(ntiles);
        std::vector<uint8_t> pad_n(C, 0);
        constexpr size_t kWave = 64;
        const int nwave = std::max(1, m->tune.block / (int)kWave);
        std::vector<int32_t> used_n(ntiles, 0);   // 64-entry positions a tile uses
        int npairs = 0, nsingles = 0;
        size_t most = 0;
        auto neg_equal = [](R a, R b) {
            const R na = -a;
            return std::memcmp(&na, &b, sizeof(R)) == 0;
        };
        // slot ids as 4*p + j during planning
        std::vector<uint32_t> pairs, sgl, bnd, sgl_h, bnd_h;
        Scalars<R> S_plan;
        S_plan.HM1 = (R)d.HM1;
        S_plan.HM2 = (R)d.HM2;
        // stable order of a task list by an integer class of each task
        std::vector<std::pair<int, uint32_t>> keyed;
        auto group_by = [&](std::vector<uint32_t> &v, auto cls) {
            keyed.resize(v.size());
            for (size_t q = 0; q < v.size(); q++) keyed[q] = {cls(v[q]), v[q]};
            std::stable_sort(keyed.begin(), keyed.end(),
                             [](const std::pair<int, uint32_t> &x, const std::pair<int, uint32_t> &y) {
                                 return x.first < y.first;
                             });
            for (size_t q = 0; q < v.size(); q++) v[q] = keyed[q].second;
        };
        for (int t = 0; t < ntiles; t++) {
            pairs.clear();
            sgl.clear();
            bnd.clear();
            sgl_h.clear();
            bnd_h.clear();
            const int p0 = t * B, pend = std::min(C, (t + 1) * B);
            for (int p = p0; p < pend; p++)
                for (int j = 0; j < 4; j++) {
                    const size_t sn = (size_t)j * C + p;
                    const int kind = meta_n[sn] & META_KIND_MASK;
                    if (kind == K_PAD) {
                        pad_n[p] |= (uint8_t)(1u << j);
                        continue;
                    }
                    const int n = nbr_n[sn];
                    // a one-slot task reading a cell of another tile
                    const bool halo = n >= 0 && (n < p0 || n >= pend);
                    const uint32_t hb = halo ? TASK_HALO : 0u;
                    if (kind >= K_B10) {
                        (halo ? bnd_h : bnd).push_back(task_word(p - p0, j, 0, 0, false, kind) | hb);
                        continue;
                    }
                    if (!halo && n >= 0 && twin_n[sn] != 0xff) {
                        const int k = twin_n[sn];
                        const size_t sk = (size_t)k * C + n;
                        const int kn = meta_n[sk] & META_KIND_MASK;
                        const bool pairable = twin_n[sk] == j && nbr_n[sk] == p &&
                                              ((kind == K_PRIMARY && kn == K_MIRROR) ||
                                               (kind == K_MIRROR && kn == K_PRIMARY)) &&
                                              neg_equal(cos_n[sn], cos_n[sk]) && neg_equal(sin_n[sn], sin_n[sk]) &&
                                              std::memcmp(&side_n[sn], &side_n[sk], sizeof(R)) == 0;
                        if (pairable) {
                            if (kind == K_PRIMARY) {
                                pairs.push_back(task_word(p - p0, j, n - p0, k, true, 0));
                                npairs++;
                            }
                            continue;
                        }
                    }
                    (halo ? sgl_h : sgl).push_back(task_word(p - p0, j, 0, 0, false, kind) | hb);
                }
            // Interior tasks by the wet/dry branches their sides take in the
            // current state (pairs: both sides wet, one side, neither; one-
            // slot tasks: wet or not), so most chunks hold one kind: a
            // wavefront runs the Riemann solve if any lane has a wet side,
            // and the dry-side ladder code if any lane has a dry one.  The
            // order of the tasks does not change the results (each slot's
            // terms have their own place; the sums run in slot order), only
            // the time; call hydro_prepare again to regroup once the wet area
            // has moved.
            auto side_wet = [&](int a, int b) {
                const R HC = std::max(h_n[b], S_plan.HM1), ZC = std::max(zbc_n[b], z_n[b]);
                return ladder_branch<R>(S_plan, h_n[a], zi_n[a], zbc_n[a], HC, zbc_n[b], ZC) == 7;
            };
            auto pair_class = [&](uint32_t w) {
                const int a = p0 + (int)task_la(w), b = p0 + (int)task_lb(w);
                return 2 - (int)side_wet(a, b) - (int)side_wet(b, a);
            };
            auto slot_class = [&](uint32_t w) {
                const int a = p0 + (int)task_la(w), n = nbr_n[(size_t)task_ja(w) * C + a];
                return (n >= 0 && side_wet(a, n)) ? 0 : 1;
            };
            group_by(pairs, pair_class);
            // leftover pairs -> one-slot tasks when the last chunk of the
            // tile's own one-slot tasks has room
            const size_t left = pairs.size() % kWave, ones = sgl.size() + bnd.size();
            const size_t room = (ones + kWave - 1) / kWave * kWave - ones;
            if (left > 0 && 2 * left <= room) {
                std::vector<uint32_t> split;
                for (size_t q = pairs.size() - left; q < pairs.size(); q++) {
                    const uint32_t w = pairs[q];
                    split.push_back(task_word(task_la(w), task_ja(w), 0, 0, false, K_PRIMARY));
                    split.push_back(task_word(task_lb(w), task_jb(w), 0, 0, false, K_MIRROR));
                }
                pairs.resize(pairs.size() - left);
                npairs -= (int)left;
                sgl.insert(sgl.begin(), split.begin(), split.end());
            }
            group_by(sgl, slot_class);
            group_by(sgl_h, slot_class);
            // Boundary slots by cost of their handler (walls and open edges
            // cheap; weirs, breaches; the fixed-point iterations of rating-
            // curve, level and discharge edges), so the costly ones share as
            // few chunks as possible: a wavefront runs every lane's path.
            auto rank = [](uint32_t w) {
                const int k = task_kind(w);
                return (k == K_B1 || k == K_B10 || k == K_B3) ? 2 : (k == K_B6 || k == K_B7) ? 1 : 0;
            };
            auto by_rank = [&](uint32_t x, uint32_t y) { return rank(x) < rank(y); };
            std::stable_sort(bnd.begin(), bnd.end(), by_rank);
            std::stable_sort(bnd_h.begin(), bnd_h.end(), by_rank);
            // one-slot tasks: the tile's own (interior, then boundary), then,
            // from a chunk of their own, those that read another tile's cells
            std::vector<uint32_t> one(sgl);
            one.insert(one.end(), bnd.begin(), bnd.end());
            const size_t nlocal = one.size(), nhalo = sgl_h.size() + bnd_h.size();
            if (nhalo > 0) one.resize((one.size() + kWave - 1) / kWave * kWave, TASK_NONE);
            one.insert(one.end(), sgl_h.begin(), sgl_h.end());
            one.insert(one.end(), bnd_h.begin(), bnd_h.end());
            nsingles += (int)(nlocal + nhalo);
            // Chunk classes: the tile's own chunks with an iterative
            // boundary handler (the longest), pairs, its other own one-slot
            // chunks, and the chunks reading other tiles (the resident
            // kernel runs those after waiting for the neighbouring tiles;
            // every other chunk needs no wait).  Placement below.
            struct Chunk {
                const std::vector<uint32_t> *src;
                size_t at, n;
                int cls;   // 0 long own, 1 pairs, 2 other own, 3 reads other tiles
            };
            std::vector<Chunk> ch;
            for (size_t q = 0; q < pairs.size(); q += kWave)
                ch.push_back({&pairs, q, std::min<size_t>(kWave, pairs.size() - q), 1});
            for (size_t q = 0; q < one.size(); q += kWave) {
                const size_t n = std::min<size_t>(kWave, one.size() - q);
                bool halo = false, heavy = false;
                for (size_t e = q; e < q + n; e++)
                    if (one[e] != TASK_NONE) {
                        halo = halo || (one[e] & TASK_HALO);
                        heavy = heavy || rank(one[e]) == 2;
                    }
                ch.push_back({&one, q, n, halo ? 3 : heavy ? 0 : 2});
            }
            std::stable_sort(ch.begin(), ch.end(), [](const Chunk &x, const Chunk &y) { return x.cls < y.cls; });
            int used = 0;
            std::vector<uint32_t> &L = lists[t];
            L.clear();
            auto put = [&](int pos, const Chunk &c) {
                used = std::max(used, pos + 1);
                if (L.size() < (size_t)(pos + 1) * kWave) L.resize((size_t)(pos + 1) * kWave, TASK_NONE);
                std::copy_n(c.src->begin() + c.at, c.n, L.begin() + (size_t)pos * kWave);
            };
            // Chunks reading other tiles first, at position 0 of the
            // wavefronts without cells (the last ones; the last wavefront
            // when every wavefront holds cells), which hold nothing else:
            // they wait for the neighbouring tiles as the step starts and run
            // while the others run the tile's own chunks.
            const int cellw = (pend - p0 + (int)kWave - 1) / (int)kWave;
            size_t nh = 0;
            while (nh < ch.size() && ch[ch.size() - 1 - nh].cls == 3) nh++;
            std::vector<int> hw;
            if (nh > 0) {
                for (int w = nwave - 1; w >= cellw; w--) hw.push_back(w);
                if (hw.empty() || nwave == 1) hw.assign(1, nwave - 1);
                if ((int)hw.size() >= nwave) hw.resize(nwave - 1);   // keep one wavefront for the rest
                if (hw.empty()) hw.push_back(0);
            }
            std::vector<char> alone(nwave, 0);   // wavefronts holding one long chunk, or the halo chunks
            for (size_t k = 0; k < nh; k++) {
                const int w = hw[k % hw.size()];
                put((int)(k / hw.size()) * nwave + w, ch[ch.size() - nh + k]);
            }
            for (int w : hw) alone[w] = (nwave > 1);
            // the tile's own chunks: position 0 of the other wavefronts (a
            // long chunk keeps its wavefront to itself), then position 1, 2,
            // ... in an order that visits the four SIMD groups in turn
            // (wavefronts w, w+4, w+8, w+12 share a SIMD), group 0 last, as
            // it holds the wavefront with the tile's last, partial row of
            // cells (or one more cell row than the others)
            const size_t nown = ch.size() - nh;
            size_t next = 0;
            for (int w = 0; w < nwave && next < nown; w++) {
                if (alone[w]) continue;
                alone[w] = ch[next].cls == 0;
                put(w, ch[next++]);
            }
            std::vector<int> order;   // wavefront order of positions 1, 2, ...
            for (int base = 0; base < nwave; base += 4)
                for (int g : {1, 2, 3, 0})
                    if (base + g < nwave && !alone[base + g]) order.push_back(base + g);
            if (order.empty())
                for (int w = 0; w < nwave; w++)
                    if (std::find(hw.begin(), hw.end(), w) == hw.end()) order.push_back(w);
            if (order.empty()) order.push_back(0);
            for (int round = 1; next < nown; round++)
                for (size_t k = 0; k < order.size() && next < nown; k++) put(round * nwave + order[k], ch[next++]);
            used_n[t] = used;
            most = std::max(most, L.size());
        }
        const size_t stride = std::max<size_t>(64, most);
        // spare entries: every thread of the last tile's workgroup loads
        // entries k0 + i and k0 + i + block unconditionally (block threads,
        // which may be more than the tile's cells)
        const size_t ntask = stride * (size_t)ntiles + 2 * (size_t)m->tune.block;
        std::vector<uint32_t> tw(ntask, TASK_NONE);
        std::vector<int32_t> tnb(ntask, -1);
        std::vector<R> tcs(2 * ntask, (R)0), tsd(ntask, (R)0);
        for (int t = 0; t < ntiles; t++) {
            const int p0 = t * B;
            for (size_t q = 0; q < lists[t].size(); q++) {
                const uint32_t w = lists[t][q];
                if (w == TASK_NONE) continue;
                const size_t k = (size_t)t * stride + q;
                const int pa = p0 + task_la(w), ja = task_ja(w);
                const size_t ea = (size_t)ja * C + pa;
                tw[k] = w;
                tnb[k] = nbr_n[ea];
                tcs[2 * k] = cos_n[ea];
                tcs[2 * k + 1] = sin_n[ea];
                tsd[k] = side_n[ea];   // a pair's two slots have the same SIDE (a pairing condition)
                // a cell another tile's one-slot task reads is a halo cell
                // (bit 4): the resident kernel publishes only those every step
                const int nb = nbr_n[ea];
                if (!(w & TASK_PAIR) && nb >= 0 && nb / B != t) pad_n[nb] |= (uint8_t)kCellHalo;
            }
        }
        // wavefronts of each tile up to its last cell that other tiles read
        // (kCellHalo): the resident kernel publishes the tile's step once
        // they have stored their cells
        std::vector<int32_t> pub(ntiles, 0);
        for (int p = 0; p < C; p++)
            if (pad_n[p] & kCellHalo) pub[p / B] = std::max(pub[p / B], (p - (p / B) * B) / (int)kWave + 1);
        if ((rc = h2d(m->tile_pub, pub))) return rc;
        m->n_pairs = npairs;
        m->n_singles = nsingles;
        m->task_stride = (int)stride;
        if ((rc = h2d(m->task_w, tw)) || (rc = h2d(m->task_nb, tnb)) || (rc = h2d(m->task_cs, tcs)) ||
            (rc = h2d(m->task_sd, tsd)) || (rc = h2d(m->task_n, used_n)) || (rc = h2d(m->cell_pad, pad_n)))
            return rc;
    }
    {
        // tiles (workgroups of `block` consecutive cells) that share an edge:
        // the persistent TILES kernel waits on exactly these between steps
        const int B = m->tile, nt = (C + B - 1) / B;
        std::vector<std::vector<int32_t>> adj(nt);
        for (int j = 0; j < 4; j++)
            for (int p = 0; p < C; p++) {
                const int q = nbr_n[(size_t)j * C + p];
                if (q < 0 || q / B == p / B) continue;
                adj[p / B].push_back(q / B);
                adj[q / B].push_back(p / B);
            }
        std::vector<int32_t> off(nt + 1, 0), lst;
        for (int t = 0; t < nt; t++) {
            std::sort(adj[t].begin(), adj[t].end());
            adj[t].erase(std::unique(adj[t].begin(), adj[t].end()), adj[t].end());
            lst.insert(lst.end(), adj[t].begin(), adj[t].end());
            off[t + 1] = (int32_t)lst.size();
        }
        m->max_tile_nb = 0;
        for (int t = 0; t < nt; t++) m->max_tile_nb = std::max(m->max_tile_nb, off[t + 1] - off[t]);
        if (lst.empty()) lst.push_back(0);
        if ((rc = h2d(m->tile_nb_off, off)) || (rc = h2d(m->tile_nb, lst)) ||
            (rc = m->tile_flag.alloc((size_t)nt * kFlagStride * sizeof(int32_t))))
            return rc;
        HIP_TRY(hipMemset(m->tile_flag.p, 0, m->tile_flag.bytes));
        m->flag_base = 0;
    }
    if ((rc = h2d(m->ZBC, zbc_n)) || (rc = h2d(m->AREA, area_n)) || (rc = h2d(m->FNC, fnc_n)) ||
        (rc = h2d(m->DTA, dta_n)) ||
        (rc = h2d(m->ZB1, zb1_n)) || (rc = h2d(m->nbr, nbr_n)) || (rc = h2d(m->meta, meta_n)) ||
        (rc = h2d(m->COSF, cos_n)) || (rc = h2d(m->SINF, sin_n)) || (rc = h2d(m->SIDE, side_n)) ||
        (rc = h2d(m->SLCOS, slc_n)) || (rc = h2d(m->SLSIN, sls_n)) || (rc = h2d(m->orig, orig_n)) ||
        (rc = h2d(m->inv, inv)) || (rc = h2d(m->twin, twin_n)))
        return rc;
    {
        // the kind of each slot's twin (the neighbour's slot facing it), for
        // the fused adjoint step's first loads; K_PAD where there is none
        std::vector<uint8_t> tk(S4, (uint8_t)K_PAD);
        for (int j = 0; j < 4; j++)
            for (int p = 0; p < C; p++) {
                const size_t sn = (size_t)j * C + p;
                const int q = nbr_n[sn];
                if (q >= 0 && twin_n[sn] <= 3) tk[sn] = meta_n[(size_t)twin_n[sn] * C + q] & META_KIND_MASK;
            }
        if ((rc = h2d(m->twin_kind, tk))) return rc;
    }
    {
        const ExtView<R> ev = ext_view<R>(m);
        if ((rc = m->ext_dev.alloc(sizeof(ExtView<R>)))) return rc;
        HIP_TRY(hipMemcpy(m->ext_dev.p, &ev, sizeof(ExtView<R>), hipMemcpyHostToDevice));
    }
    for (int b = 0; b < 2; b++)
        if ((rc = m->st[b].alloc((size_t)C * sizeof(State<R>)))) return rc;
    if (!m->tick.p && (rc = m->tick.alloc(sizeof(StepClock)))) return rc;
    if (!m->barrier.p && (rc = m->barrier.alloc(sizeof(unsigned)))) return rc;
    if (!m->perr.p && (rc = m->perr.alloc(sizeof(int)))) return rc;
    HIP_TRY(hipMemset(m->perr.p, 0, sizeof(int)));
    m->graphs.clear();
    m->prepared = true;
    m->since_group = 0;
    m->int_valid = false;
    return HYDRO_OK;
}

template <class R> int ensure_internal(hydro_mesh *m, hipStream_t s) {
    if (!m->prepared) {
        const hydro_tune t = m->tune;
        int rc = hydro_prepare(m, &t);
        if (rc) return rc;
    }
    if (m->int_valid) return HYDRO_OK;
    const ExtView<R> ev = ext_view<R>(m);
    m->cur = 0;
    HIP_TRY(launch_gather<R>(ev, m->orig.as<int32_t>(), m->st[0].as<State<R>>(), m->CELL, s));
    HIP_TRY(launch_refresh_meta((const int32_t *)m->d.KLAS, m->orig.as<int32_t>(), m->meta.as<uint8_t>(), m->CELL,
                                s));
    m->int_valid = true;
    return HYDRO_OK;
}

// A persistent launch whose waits exceeded their spin limit left early with
// a partial state: report it at the next call that reads the state or
// launches again (one synchronous 4-byte read per persistent launch).
int check_persistent(hydro_mesh *m, hipStream_t s) {
    if (!m->persist_pending || m->defer_check) return HYDRO_OK;
    m->persist_pending = false;
    int err = 0;
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipMemcpy(&err, m->perr.p, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
        HIP_TRY(hipMemset(m->perr.p, 0, sizeof(int)));
        // the tiles stopped at different steps: start the flags afresh
        if (m->tile_flag.p) HIP_TRY(hipMemset(m->tile_flag.p, 0, m->tile_flag.bytes));
        m->flag_base = 0;
        return fail(HYDRO_ERR_HIP, "a persistent launch stopped: a wait between steps exceeded its spin limit");
    }
    return HYDRO_OK;
}

template <class R> int ensure_external(hydro_mesh *m, hipStream_t s) {
    int rc = check_persistent(m, s);
    if (rc) return rc;
    if (m->ext_valid) return HYDRO_OK;
    const ExtView<R> ev = ext_view<R>(m);
    HIP_TRY(launch_scatter<R>(ev, m->orig.as<int32_t>(), m->st[m->cur].as<State<R>>(), m->CELL, s));
    m->ext_valid = true;
    return HYDRO_OK;
}

void advance(hydro_mesh *m) {
    m->kt++;
    if (m->kt >= m->spd) {
        m->kt = 1;
        m->day++;
    }
}

int64_t global_step(const hydro_mesh *m) { return (int64_t)m->day * (m->spd - 1) + (m->kt - 1); }

template <class R, int VAR> int steps_two_kernel(hydro_mesh *m, int n, hipStream_t s) {
    int rc = ensure_external<R>(m, s);
    if (rc) return rc;
    const ExtView<R> ev = ext_view<R>(m);
    for (int i = 0; i < n; i++) {
        HIP_TRY((launch_calculate_flux<R, VAR>(ev, m->kt, m->day, m->block_two, s)));
        HIP_TRY((launch_update_cell<R, VAR>(ev, m->block_two, s)));
        advance(m);
    }
    m->int_valid = false;
    m->last_launches += 2 * n;
    return HYDRO_OK;
}

template <class R, int VAR> int steps_fused(hydro_mesh *m, int n, hipStream_t s) {
    int rc = ensure_internal<R>(m, s);
    if (rc) return rc;
    for (int i = 0; i < n; i++) {
        const FusedView<R> f = fused_view<R>(m, m->cur);
        HIP_TRY((launch_fused_step<R, VAR>(f, m->kt, m->day, m->tune.block, m->tune.xcd_remap, s)));
        m->cur ^= 1;
        advance(m);
    }
    m->ext_valid = false;
    m->last_launches += n;
    return HYDRO_OK;
}

__global__ void k_set_tick(int64_t *tick, int64_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *tick = v;
}

__global__ void k_zero_u32(unsigned *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0u;
}

template <class R, int VAR> int build_graph(hydro_mesh *m, int parity, int chunk, GraphExec &ge) {
    hipStream_t cs = m->cap_stream;
    HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    hipError_t err = hipSuccess;
    int cur = parity;
    for (int k = 0; k < chunk && err == hipSuccess; k++) {
        const FusedView<R> f = fused_view<R>(m, cur);
        err = launch_fused_step_tick<R, VAR>(f, m->tick.as<StepClock>(), k, m->spd, m->tune.block,
                                             m->tune.xcd_remap, cs);
        cur ^= 1;
    }
    if (err == hipSuccess) err = launch_clock_advance(m->tick.as<StepClock>(), chunk, m->spd, cs);
    hipError_t e2 = hipStreamEndCapture(cs, &ge.g);
    if (err != hipSuccess) return fail(HYDRO_ERR_HIP, "graph capture launch failed: %s", hipGetErrorString(err));
    if (e2 != hipSuccess) return fail(HYDRO_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e2));
    HIP_TRY(hipGraphInstantiate(&ge.x, ge.g, nullptr, nullptr, 0));
    return HYDRO_OK;
}

template <class R, int VAR> int steps_graph(hydro_mesh *m, int n, hipStream_t s) {
    int rc = ensure_internal<R>(m, s);
    if (rc) return rc;
    int chunk = m->tune.graph_chunk;
    if (chunk < 2) chunk = 2;
    if (chunk & 1) chunk++;   // even, so buffer parity is unchanged by one replay
    if (!m->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&m->cap_stream, hipStreamNonBlocking));
    int done = 0;
    if (n >= chunk) {
        GraphKey key{HYDRO_MODE_GRAPH, m->cur, chunk};
        auto it = m->graphs.find(key);
        if (it == m->graphs.end()) {
            auto ge = std::make_unique<GraphExec>();
            if ((rc = build_graph<R, VAR>(m, m->cur, chunk, *ge))) return rc;
            it = m->graphs.emplace(key, std::move(ge)).first;
        }
        HIP_TRY(launch_clock_set(m->tick.as<StepClock>(), m->day, m->kt, s));
        while (n - done >= chunk) {
            HIP_TRY(hipGraphLaunch(it->second->x, s));
            for (int k = 0; k < chunk; k++) advance(m);
            done += chunk;
            m->last_launches += chunk;
        }
    }
    if (done < n) return steps_fused<R, VAR>(m, n - done, s);
    m->ext_valid = false;
    return HYDRO_OK;
}

// TWO_KERNEL steps captured into a hipGraph: `chunk` pairs of flux and
// update launches on the reference layout, the flux kernels reading the
// step's (day, kt) from the device clock, then the clock advance, so every
// replay moves the boundary series on.  The layout's buffers do not
// alternate, so one graph serves every replay.
template <class R, int VAR> int build_graph_two_kernel(hydro_mesh *m, int chunk, GraphExec &ge) {
    hipStream_t cs = m->cap_stream;
    const ExtView<R> ev = ext_view<R>(m);
    HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    hipError_t err = hipSuccess;
    for (int k = 0; k < chunk && err == hipSuccess; k++) {
        err = launch_calculate_flux_tick<R, VAR>(ev, m->tick.as<StepClock>(), k, m->spd, m->block_two, cs);
        if (err == hipSuccess) err = launch_update_cell<R, VAR>(ev, m->block_two, cs);
    }
    if (err == hipSuccess) err = launch_clock_advance(m->tick.as<StepClock>(), chunk, m->spd, cs);
    hipError_t e2 = hipStreamEndCapture(cs, &ge.g);
    if (err != hipSuccess) return fail(HYDRO_ERR_HIP, "graph capture launch failed: %s", hipGetErrorString(err));
    if (e2 != hipSuccess) return fail(HYDRO_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e2));
    HIP_TRY(hipGraphInstantiate(&ge.x, ge.g, nullptr, nullptr, 0));
    return HYDRO_OK;
}

template <class R, int VAR> int steps_graph_two_kernel(hydro_mesh *m, int n, hipStream_t s) {
    int rc = ensure_external<R>(m, s);
    if (rc) return rc;
    if (!m->tick.p && (rc = m->tick.alloc(sizeof(StepClock)))) return rc;
    const int chunk = std::max(1, m->tune.graph_chunk);
    if (!m->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&m->cap_stream, hipStreamNonBlocking));
    int done = 0;
    if (n >= chunk) {
        GraphKey key{HYDRO_MODE_GRAPH_TWO_KERNEL, 0, chunk};
        auto it = m->graphs.find(key);
        if (it == m->graphs.end()) {
            auto ge = std::make_unique<GraphExec>();
            if ((rc = build_graph_two_kernel<R, VAR>(m, chunk, *ge))) return rc;
            it = m->graphs.emplace(key, std::move(ge)).first;
        }
        HIP_TRY(launch_clock_set(m->tick.as<StepClock>(), m->day, m->kt, s));
        while (n - done >= chunk) {
            HIP_TRY(hipGraphLaunch(it->second->x, s));
            for (int k = 0; k < chunk; k++) advance(m);
            done += chunk;
            m->last_launches += 2 * chunk;
        }
        m->int_valid = false;
    }
    if (done < n) return steps_two_kernel<R, VAR>(m, n - done, s);
    return HYDRO_OK;
}

template <class R, int VAR> int steps_persistent(hydro_mesh *m, int n, hipStream_t s) {
    int rc = check_persistent(m, s);
    if (rc || (rc = ensure_internal<R>(m, s))) return rc;
    if (n == 0) return HYDRO_OK;
    const int block = m->tune.block;
    const int ntiles = (m->CELL + m->tile - 1) / m->tile;
    const int want = m->tune.persistent_wgs_per_cu;
    const FusedView<R> f = fused_view<R>(m, m->cur);
    PersistentArgs pa;
    pa.nsteps = n;
    pa.day0 = m->day;
    pa.kt0 = m->kt;
    pa.spd = m->spd;
    pa.barrier = m->barrier.as<unsigned>();
    pa.error = m->perr.as<int>();
    pa.spin_limit = m->tune.spin_limit > 0 ? (unsigned)m->tune.spin_limit : 1u << 24;
    pa.flag0 = 0;
    // the resident kernel when every tile fits on the GPU at once (the
    // capacity is asked once per prepared layout: nothing but the launch
    // itself is left on the per-call path)
    bool resident = false;
    if (f.scheme == HYDRO_SCHEME_TILES && !m->tune.persistent_loop && want == 0) {
        if (m->resident_cap < 0) HIP_TRY((resident_capacity<R, VAR>(block, m->rare, &m->resident_cap)));
        resident = m->resident_cap >= f.ntiles;
    }
    if (resident) {
        // the flags need no reset: this launch counts on from where the
        // last one left them (restarting from 0 long before int32 overflow)
        if ((int64_t)m->flag_base + n > (1 << 30)) {
            HIP_TRY(launch_fill_i32(m->tile_flag.as<int32_t>(), 0, ntiles * kFlagStride, s));
            m->flag_base = 0;
        }
        pa.flag0 = m->flag_base;
        HIP_TRY((launch_resident<R, VAR>(f, pa, m->tune.xcd_remap, m->tune.resident_launch == 1, s)));
        m->flag_base += n;
    } else {
        int maxg = 0;
        HIP_TRY((persistent_max_grid<R, VAR>(block, m->rare, m->tune.scheme, m->tune.waves_per_simd == 4, &maxg)));
        int grid = maxg;
        if (want > 0) {
            int dev = 0, cus = 0;
            HIP_TRY(hipGetDevice(&dev));
            HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            grid = std::min(maxg, want * cus);
        }
        grid = std::max(1, std::min(grid, ntiles));
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, pa.barrier);
        HIP_TRY(hipGetLastError());
        if (f.scheme == HYDRO_SCHEME_TILES) HIP_TRY(launch_fill_i32(m->tile_flag.as<int32_t>(), 0, ntiles * kFlagStride, s));
        HIP_TRY((launch_persistent<R, VAR>(f, pa, block, grid, s)));
        m->flag_base = (f.scheme == HYDRO_SCHEME_TILES) ? n : m->flag_base;
    }
    m->last_resident = resident;
    m->persist_pending = true;
    if (n & 1) m->cur ^= 1;
    for (int i = 0; i < n; i++) advance(m);
    m->ext_valid = false;
    m->last_launches += 1;
    return HYDRO_OK;
}

template <class R, int VAR> int step_dispatch(hydro_mesh *m, int n, int mode, hipStream_t s) {
    switch (mode) {
    case HYDRO_MODE_TWO_KERNEL: return steps_two_kernel<R, VAR>(m, n, s);
    case HYDRO_MODE_FUSED: return steps_fused<R, VAR>(m, n, s);
    case HYDRO_MODE_GRAPH: return steps_graph<R, VAR>(m, n, s);
    case HYDRO_MODE_PERSISTENT: return steps_persistent<R, VAR>(m, n, s);
    case HYDRO_MODE_GRAPH_TWO_KERNEL: return steps_graph_two_kernel<R, VAR>(m, n, s);
    default: return fail(HYDRO_ERR_UNSUPPORTED, "mode %d not supported", mode);
    }
}

hydro_tune default_tune() {
    hydro_tune t{};
    t.block = 256;
    t.reorder = HYDRO_REORDER_TILES;
    t.xcd_remap = 1;
    t.graph_chunk = 16;
    t.persistent_wgs_per_cu = 0;
    t.materialize_flux = 0;
    t.waves_per_simd = 0;
    t.scheme = HYDRO_SCHEME_TILES;
    t.tile_cells = 0;
    t.spin_limit = 0;
    t.persistent_loop = 0;
    t.resident_launch = 0;
    t.regroup_steps = 0;
    return t;
}

// An index array (NAC or KLAS) of n entries as int32: read as it is, or as
// Real values of the mesh's dtype that must be whole numbers in int32 range.
int read_index(std::vector<int32_t> &out, const void *p, size_t n, bool real, int dtype, const char *name) {
    if (!real) return d2h(out, p, n);
    std::vector<double> v(n);
    if (dtype == HYDRO_F64) {
        int rc = d2h(v, p, n);
        if (rc) return rc;
    } else {
        std::vector<float> f;
        int rc = d2h(f, p, n);
        if (rc) return rc;
        for (size_t i = 0; i < n; i++) v[i] = f[i];
    }
    out.resize(n);
    for (size_t i = 0; i < n; i++) {
        if (!(v[i] >= -2147483648.0 && v[i] <= 2147483647.0) || v[i] != std::floor(v[i]))
            return fail(HYDRO_ERR_INVALID, "%s[%zu] = %g is not a whole number", name, i, v[i]);
        out[i] = (int32_t)v[i];
    }
    return HYDRO_OK;
}

int check_ptr(const void *p, const char *name, bool required) {
    if (!p) return required ? fail(HYDRO_ERR_INVALID, "%s is NULL", name) : HYDRO_OK;
    if (!is_device_ptr(p)) return fail(HYDRO_ERR_INVALID, "%s is not a device pointer", name);
    return HYDRO_OK;
}

}  // namespace

// Regroup the pair tasks by the current state (on the device, stream-ordered).
// (A persistent launch that stopped at its spin limit is reported by the
// next call that steps or reads the state; regrouping a partial state is
// harmless, so this call does not wait for that check.)
template <class R> int regroup_t(hydro_mesh *m, hipStream_t s) {
    const int rc = ensure_internal<R>(m, s);
    if (rc) return rc;
    if (m->tune.scheme != HYDRO_SCHEME_TILES) return HYDRO_OK;
    HIP_TRY(launch_regroup<R>(fused_view<R>(m, m->cur), s));
    m->since_group = 0;
    return HYDRO_OK;
}

// Pair chunks of the plan, and those holding pairs of more than one wet/dry
// class in the current state (host-side count, synchronous).
template <class R> int plan_grouping_t(hydro_mesh *m, int *pair_chunks, int *mixed_chunks, hipStream_t s) {
    int rc = check_persistent(m, s);
    if (rc || (rc = ensure_internal<R>(m, s))) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    *pair_chunks = *mixed_chunks = 0;
    if (m->tune.scheme != HYDRO_SCHEME_TILES) return HYDRO_OK;
    const int C = m->CELL, B = m->tile, nt = (C + B - 1) / B;
    std::vector<uint32_t> tw;
    std::vector<int32_t> tn;
    std::vector<State<R>> st;
    std::vector<R> zbc, zb1;
    if ((rc = d2h(tw, m->task_w.p, m->task_w.bytes / sizeof(uint32_t))) ||
        (rc = d2h(tn, m->task_n.p, (size_t)nt)) || (rc = d2h(st, m->st[m->cur].p, (size_t)C)) ||
        (rc = d2h(zbc, m->ZBC.p, (size_t)C)) || (m->need_zb1 && (rc = d2h(zb1, m->ZB1.p, (size_t)C))))
        return rc;
    Scalars<R> S;
    S.HM1 = (R)m->d.HM1;
    S.HM2 = (R)m->d.HM2;
    auto wet = [&](int a, int b) {
        const R zb1a = m->need_zb1 ? zb1[a] : (R)(zbc[a] + S.HM1);
        return ladder_branch<R>(S, st[a].H, std::max(st[a].Z, zb1a), zbc[a], std::max(st[b].H, S.HM1), zbc[b],
                                std::max(zbc[b], st[b].Z)) == 7;
    };
    for (int t = 0; t < nt; t++)
        for (int q = 0; q < tn[t]; q++) {
            int seen = 0, pairs = 0, others = 0;
            for (int l = 0; l < 64; l++) {
                const uint32_t w = tw[(size_t)t * m->task_stride + (size_t)q * 64 + l];
                if (w == TASK_NONE) continue;
                if (!(w & TASK_PAIR)) {
                    others++;
                    continue;
                }
                const int a = t * B + (int)task_la(w), b = t * B + (int)task_lb(w);
                seen |= 1 << (2 - (int)wet(a, b) - (int)wet(b, a));
                pairs++;
            }
            if (pairs == 0 || others > 0) continue;
            (*pair_chunks)++;
            if (seen & (seen - 1)) (*mixed_chunks)++;
        }
    return HYDRO_OK;
}

// The reordered modes of the TILES scheme in launches of at most R steps
// (hydro_tune.regroup_steps, 0 = kRegroupSteps, -1 = no limit), the plan's
// tasks regrouped by the current state before a launch once R steps have
// run since the last grouping.
constexpr int kRegroupSteps = 200;
// Steps between regroupings for loops that launch step by step (0 = none).
int regroup_every(const hydro_mesh *m) {
    if (m->tune.scheme != HYDRO_SCHEME_TILES || m->tune.regroup_steps < 0) return 0;
    return m->tune.regroup_steps == 0 ? kRegroupSteps : m->tune.regroup_steps;
}
template <class R, int VAR> int steps_grouped(hydro_mesh *m, int n, int mode, hipStream_t s) {
    int every = m->tune.regroup_steps == 0 ? kRegroupSteps : m->tune.regroup_steps;
    const bool reordered = mode == HYDRO_MODE_FUSED || mode == HYDRO_MODE_GRAPH || mode == HYDRO_MODE_PERSISTENT;
    if (mode == HYDRO_MODE_GRAPH && every > 0) {   // whole replays of steps_graph's chunk per stretch
        const int chunk = std::max(2, m->tune.graph_chunk + (m->tune.graph_chunk & 1));
        every = std::max(chunk, every / chunk * chunk);
    }
    if (every < 0 || !reordered || m->tune.scheme != HYDRO_SCHEME_TILES) return step_dispatch<R, VAR>(m, n, mode, s);
    int rc = HYDRO_OK;
    if (n == 0) return step_dispatch<R, VAR>(m, 0, mode, s);
    // the first launch checks (synchronously) whether the previous persistent
    // launch stopped early; the later ones of this request do not wait for
    // each other: a stop is reported by the next call that steps or reads
    bool first = true;
    while (n > 0 && rc == HYDRO_OK) {
        m->defer_check = !first;
        if (m->prepared && m->since_group >= every) rc = regroup_t<R>(m, s);
        const int seg = std::min(n, every);
        if (rc == HYDRO_OK) rc = step_dispatch<R, VAR>(m, seg, mode, s);
        m->since_group += seg;
        n -= seg;
        first = false;
    }
    m->defer_check = false;
    return rc;
}

extern "C" {

const char *hydro_last_error(void) { return g_err.c_str(); }
int hydro_abi_version(void) { return HYDRO_ABI_VERSION; }

int hydro_mesh_create(const hydro_mesh_desc *desc, hydro_mesh **out) {
    if (!desc || !out) return fail(HYDRO_ERR_INVALID, "null argument");
    *out = nullptr;
    const hydro_mesh_desc &d = *desc;
    if (d.abi_version != HYDRO_ABI_VERSION)
        return fail(HYDRO_ERR_INVALID, "abi_version %d, library is %d", d.abi_version, HYDRO_ABI_VERSION);
    if (d.dtype != HYDRO_F64 && d.dtype != HYDRO_F32) return fail(HYDRO_ERR_INVALID, "dtype %d", d.dtype);
    if (d.variant != HYDRO_VARIANT_A && d.variant != HYDRO_VARIANT_B)
        return fail(HYDRO_ERR_INVALID, "variant %d", d.variant);
    if (d.CELL < 0 || d.CELL > (1 << 28)) return fail(HYDRO_ERR_INVALID, "CELL %d", d.CELL);
    if (d.NDAYS < 1) return fail(HYDRO_ERR_INVALID, "NDAYS %d", d.NDAYS);
    if (d.NHQ < 0) return fail(HYDRO_ERR_INVALID, "NHQ %d", d.NHQ);
    if (d.index_real & ~(HYDRO_NAC_REAL | HYDRO_KLAS_REAL)) return fail(HYDRO_ERR_INVALID, "index_real %d", d.index_real);
    if (!(d.DT > 0) || d.MDT <= 0) return fail(HYDRO_ERR_INVALID, "DT %g MDT %d", d.DT, d.MDT);
    if (!(d.HM1 > 0) || !(d.HM2 >= d.HM1)) return fail(HYDRO_ERR_INVALID, "HM1 %g HM2 %g", d.HM1, d.HM2);
    int rc;
    const bool need = d.CELL > 0;
#define CHK(ptr, req) \
    if ((rc = check_ptr(d.ptr, #ptr, (req)))) return rc
    CHK(H, need); CHK(U, need); CHK(V, need); CHK(Z, need); CHK(W, need);
    CHK(ZBC, need); CHK(ZB1, need); CHK(AREA, need); CHK(FNC, need);
    CHK(NAC, need); CHK(KLAS, need); CHK(SIDE, need); CHK(COSF, need); CHK(SINF, need);
    CHK(SLCOS, need); CHK(SLSIN, need);
    CHK(FLUX0, need); CHK(FLUX1, need); CHK(FLUX2, need); CHK(FLUX3, need);
    CHK(BoundaryFeature, false); CHK(NHQ1, false); CHK(ZW, false); CHK(QW, false);
    CHK(QT, false); CHK(DQT, false); CHK(ZT, false); CHK(DZT, false);
#undef CHK
    auto m = std::make_unique<hydro_mesh>();
    m->d = d;
    m->CELL = d.CELL;
    m->rsize = d.dtype == HYDRO_F64 ? 8 : 4;
    m->spd = (int)(d.MDT / d.DT);
    m->tune = default_tune();
    // Host-side validation of the index arrays the kernels dereference
    // (Real-typed ones are converted to owned int32 copies here).
    const size_t S4 = 4 * (size_t)d.CELL;
    std::vector<int32_t> NAC, KLAS;
    if ((rc = read_index(NAC, d.NAC, S4, (d.index_real & HYDRO_NAC_REAL) != 0, d.dtype, "NAC")) ||
        (rc = read_index(KLAS, d.KLAS, S4, (d.index_real & HYDRO_KLAS_REAL) != 0, d.dtype, "KLAS")))
        return rc;
    if (d.index_real & HYDRO_NAC_REAL) {
        if ((rc = h2d(m->nac_own, NAC))) return rc;
        m->d.NAC = m->nac_own.p;
    }
    if (d.index_real & HYDRO_KLAS_REAL) {
        if ((rc = h2d(m->klas_own, KLAS))) return rc;
        m->klas_real = d.KLAS;
        m->d.KLAS = m->klas_own.p;
    }
    for (size_t s = 0; s < S4; s++) {
        if (NAC[s] < 0 || NAC[s] > d.CELL)
            return fail(HYDRO_ERR_INVALID, "NAC[%zu] = %d outside [0, CELL=%d]", s, NAC[s], d.CELL);
        m->has_kind[kind_of_klas(KLAS[s], false)] = true;
    }
    m->rare = m->has_kind[K_B3] || m->has_kind[K_B5] || m->has_kind[K_B6] || m->has_kind[K_B7] ||
              m->has_kind[K_BUNK];
    if (m->has_kind[K_B10] && (!d.QT || !d.DQT)) return fail(HYDRO_ERR_INVALID, "KLAS 10 slots need QT and DQT");
    if (m->has_kind[K_B1] && (!d.ZT || !d.DZT)) return fail(HYDRO_ERR_INVALID, "KLAS 1 slots need ZT and DZT");
    if (m->has_kind[K_B3] && (!d.ZW || !d.QW || !d.NHQ1 || d.NHQ < 2))
        return fail(HYDRO_ERR_INVALID, "KLAS 3 slots need ZW, QW, NHQ1 and NHQ >= 2");
    if ((m->has_kind[K_B6] || m->has_kind[K_B7]) && !d.BoundaryFeature)
        return fail(HYDRO_ERR_INVALID, "KLAS 6/7 slots need BoundaryFeature");
    if (m->has_kind[K_B3]) {
        std::vector<int32_t> nhq1;
        if ((rc = d2h(nhq1, d.NHQ1, d.CELL))) return rc;
        for (int i = 0; i < d.CELL; i++)
            if (nhq1[i] < 0 || nhq1[i] > d.NHQ)
                return fail(HYDRO_ERR_INVALID, "NHQ1[%d] = %d outside [0, NHQ]", i, nhq1[i]);
    }
    *out = m.release();
    return HYDRO_OK;
}

int hydro_mesh_destroy(hydro_mesh *m) {
    delete m;
    return HYDRO_OK;
}

int hydro_prepare(hydro_mesh *m, const hydro_tune *tune) {
    if (!m) return fail(HYDRO_ERR_INVALID, "null mesh");
    hydro_tune t = tune ? *tune : default_tune();
    {
        const int most = (t.scheme == HYDRO_SCHEME_TILES) ? kTileMaxBlock : kFusedMaxBlock;
        if (t.block < 64 || t.block > most || (t.block % 64))
            return fail(HYDRO_ERR_INVALID, "block %d (64..%d, multiple of 64, for scheme %d)", t.block, most,
                        t.scheme);
    }
    if (t.reorder < 0 || t.reorder > 3) return fail(HYDRO_ERR_INVALID, "reorder %d", t.reorder);
    if (t.reorder == HYDRO_REORDER_HILBERT && (int)m->xc.size() != m->CELL)
        return fail(HYDRO_ERR_UNSUPPORTED, "the Hilbert order needs cell coordinates: call hydro_set_cell_coords");
    if (t.waves_per_simd != 0 && !(t.waves_per_simd == 4 && t.scheme == HYDRO_SCHEME_TILES))
        return fail(HYDRO_ERR_INVALID, "waves_per_simd %d: 0, or 4 with the TILES scheme", t.waves_per_simd);
    if (t.scheme != HYDRO_SCHEME_TILES && t.scheme != HYDRO_SCHEME_CELL)
        return fail(HYDRO_ERR_INVALID, "scheme %d", t.scheme);
    if (t.tile_cells != 0 && t.tile_cells != HYDRO_TILE_BALANCED && (t.tile_cells < 64 || t.tile_cells > t.block))
        return fail(HYDRO_ERR_INVALID, "tile_cells %d (0, -1, or 64..block)", t.tile_cells);
    if (t.tile_cells == HYDRO_TILE_BALANCED && t.scheme != HYDRO_SCHEME_TILES)
        return fail(HYDRO_ERR_INVALID, "tile_cells -1 (balanced) needs the TILES scheme");
    if (t.spin_limit < 0) return fail(HYDRO_ERR_INVALID, "spin_limit %d", t.spin_limit);
    if (t.persistent_loop != 0 && t.persistent_loop != 1)
        return fail(HYDRO_ERR_INVALID, "persistent_loop %d (0 or 1)", t.persistent_loop);
    if (t.resident_launch != 0 && t.resident_launch != 1)
        return fail(HYDRO_ERR_INVALID, "resident_launch %d (0 or 1)", t.resident_launch);
    if (t.regroup_steps < -1) return fail(HYDRO_ERR_INVALID, "regroup_steps %d (-1, 0 or more)", t.regroup_steps);
    int tile = t.block;
    if (t.scheme == HYDRO_SCHEME_TILES && t.tile_cells > 0) {
        tile = t.tile_cells;
    } else if (t.tile_cells == HYDRO_TILE_BALANCED) {
        // as many tiles as the GPU holds workgroups of `block` threads at
        // once (16 waves per CU), every CU the same number, so no CU carries
        // one tile more than the others; at most `block` cells per tile
        int dev = 0, cus = 0;
        HIP_TRY(hipGetDevice(&dev));
        HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const int64_t slots = (int64_t)std::max(1, cus) * std::max(1, kTileMaxBlock / t.block);
        tile = (int)std::min<int64_t>(t.block, std::max<int64_t>(64, ((int64_t)m->CELL + slots - 1) / slots));
    } else if (t.scheme == HYDRO_SCHEME_TILES) {
        // a mesh with fewer full tiles than CUs leaves CUs idle: cut it into
        // one tile per CU instead (a tile's time is mostly the latency of its
        // edge solves, which smaller tiles shorten)
        int dev = 0, cus = 0;
        HIP_TRY(hipGetDevice(&dev));
        HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const int C = m->CELL;
        if (cus > 0 && C > 0 && ((int64_t)C + t.block - 1) / t.block < cus)
            tile = (int)std::max<int64_t>(64, ((int64_t)C + cus - 1) / cus);
    }
    // the internal state must be written back before the layout is replaced
    int rc = HYDRO_OK;
    if (m->prepared && !m->ext_valid)
        rc = (m->d.dtype == HYDRO_F64) ? ensure_external<double>(m, nullptr) : ensure_external<float>(m, nullptr);
    if (rc) return rc;
    HIP_TRY(hipDeviceSynchronize());
    m->tune = t;
    m->resident_cap = -1;   // block and rare may have changed
    m->tile = tile;
    m->prepared = false;
    return (m->d.dtype == HYDRO_F64) ? prepare_t<double>(m) : prepare_t<float>(m);
}

int hydro_calculate_flux(hydro_mesh *m, int kt, int jt, hipStream_t stream) {
    if (!m) return fail(HYDRO_ERR_INVALID, "null mesh");
    if (jt < 0 || jt >= m->d.NDAYS) return fail(HYDRO_ERR_RANGE, "jt %d outside [0, NDAYS=%d)", jt, m->d.NDAYS);
    int rc;
    if (m->d.dtype == HYDRO_F64) {
        if ((rc = ensure_external<double>(m, stream))) return rc;
        const ExtView<double> ev = ext_view<double>(m);
        HIP_TRY(m->d.variant ? launch_calculate_flux<double, 1>(ev, kt, jt, m->block_two, stream)
                             : launch_calculate_flux<double, 0>(ev, kt, jt, m->block_two, stream));
    } else {
        if ((rc = ensure_external<float>(m, stream))) return rc;
        const ExtView<float> ev = ext_view<float>(m);
        HIP_TRY(m->d.variant ? launch_calculate_flux<float, 1>(ev, kt, jt, m->block_two, stream)
                             : launch_calculate_flux<float, 0>(ev, kt, jt, m->block_two, stream));
    }
    m->int_valid = false;
    return HYDRO_OK;
}

int hydro_update_cell(hydro_mesh *m, hipStream_t stream) {
    if (!m) return fail(HYDRO_ERR_INVALID, "null mesh");
    int rc;
    if (m->d.dtype == HYDRO_F64) {
        if ((rc = ensure_external<double>(m, stream))) return rc;
        const ExtView<double> ev = ext_view<double>(m);
        HIP_TRY(m->d.variant ? launch_update_cell<double, 1>(ev, m->block_two, stream)
                             : launch_update_cell<double, 0>(ev, m->block_two, stream));
    } else {
        if ((rc = ensure_external<float>(m, stream))) return rc;
        const ExtView<float> ev = ext_view<float>(m);
        HIP_TRY(m->d.variant ? launch_update_cell<float, 1>(ev, m->block_two, stream)
                             : launch_update_cell<float, 0>(ev, m->block_two, stream));
    }
    m->int_valid = false;
    return HYDRO_OK;
}

int hydro_step_n(hydro_mesh *m, int nsteps, int mode, hipStream_t stream) {
    if (!m) return fail(HYDRO_ERR_INVALID, "null mesh");
    if (nsteps < 0) return fail(HYDRO_ERR_INVALID, "nsteps %d", nsteps);
    const int per = m->spd - 1;
    const int64_t total = per > 0 ? (int64_t)m->d.NDAYS * per : 0;
    const int64_t remaining = total - (per > 0 ? global_step(m) : 0);
    if (nsteps > remaining)
        return fail(HYDRO_ERR_RANGE, "%d steps requested, %lld remain before the NDAYS cap", nsteps,
                    (long long)remaining);
    m->last_launches = 0;
    m->last_steps = nsteps;
    if (m->profiling) {
        if (!m->ev0) HIP_TRY(hipEventCreate(&m->ev0));
        if (!m->ev1) HIP_TRY(hipEventCreate(&m->ev1));
        HIP_TRY(hipEventRecord(m->ev0, stream));
    }
    int rc;
    if (m->d.dtype == HYDRO_F64)
        rc = m->d.variant ? steps_grouped<double, 1>(m, nsteps, mode, stream)
                          : steps_grouped<double, 0>(m, nsteps, mode, stream);
    else
        rc = m->d.variant ? steps_grouped<float, 1>(m, nsteps, mode, stream)
                          : steps_grouped<float, 0>(m, nsteps, mode, stream);
    if (rc) return rc;
    if (m->profiling) {
        HIP_TRY(hipEventRecord(m->ev1, stream));
        HIP_TRY(hipEventSynchronize(m->ev1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, m->ev0, m->ev1));
        m->last_ms = ms;
    }
    return HYDRO_OK;
}

int hydro_set_time(hydro_mesh *m, int day, int kt) {
    if (!m) return fail(HYDRO_ERR_INVALID, "null mesh");
    if (day < 0 || day > m->d.NDAYS || kt < 1 || (kt >= m->spd && m->spd > 1))
        return fail(HYDRO_ERR_RANGE, "time (day %d, kt %d) outside the step grid", day, kt);
    m->day = day;
    m->kt = kt;
    return HYDRO_OK;
}

int hydro_get_time(const hydro_mesh *m, int *day, int *kt, int *steps_per_day) {
    if (!m) return fail(HYDRO_ERR_INVALID, "null mesh");
    if (day) *day = m->day;
    if (kt) *kt = m->kt;
    if (steps_per_day) *steps_per_day = m->spd;
    return HYDRO_OK;
}

int hydro_sync(hydro_mesh *m, hipStream_t stream) {
    if (!m) return fail(HYDRO_ERR_INVALID, "null mesh");
    return (m->d.dtype == HYDRO_F64) ? ensure_external<double>(m, stream) : ensure_external<float>(m, stream);
}

// ---- autotuner ---------------------------------------------------------------

namespace {
// Copies of the caller's mutable arrays (state, W, KLAS, FLUX0..3), so every
// candidate starts from the same step.
struct Snapshot {
    DevBuf buf[11];
    void *src[11] = {};
    size_t bytes[11] = {};
    int n = 0;
    int take(hydro_mesh *m, hipStream_t s) {
        const size_t rs = m->d.dtype == HYDRO_F64 ? 8 : 4, C = (size_t)m->CELL;
        void *p[11] = {m->d.H, m->d.U, m->d.V, m->d.Z, m->d.W, m->d.KLAS, m->d.FLUX0, m->d.FLUX1, m->d.FLUX2,
                       m->d.FLUX3, m->klas_real};
        const size_t b[11] = {C * rs, C * rs, C * rs, C * rs, C * rs, 4 * C * 4, 4 * C * rs, 4 * C * rs, 4 * C * rs,
                              4 * C * rs, 4 * C * rs};
        for (int i = 0; i < 11; i++) {
            if (!p[i]) continue;
            int rc = buf[n].alloc(b[i]);
            if (rc) return rc;
            src[n] = p[i];
            bytes[n] = b[i];
            HIP_TRY(hipMemcpyAsync(buf[n].p, p[i], b[i], hipMemcpyDeviceToDevice, s));
            n++;
        }
        return HYDRO_OK;
    }
    int put(hipStream_t s) const {
        for (int i = 0; i < n; i++) HIP_TRY(hipMemcpyAsync(src[i], buf[i].p, bytes[i], hipMemcpyDeviceToDevice, s));
        return HYDRO_OK;
    }
};

// Back to the snapshot's state at (day, kt), in the mesh's current layout.
int restore_snapshot(hydro_mesh *m, const Snapshot &snap, int day, int kt, hipStream_t s) {
    int rc = (m->d.dtype == HYDRO_F64) ? ensure_external<double>(m, s) : ensure_external<float>(m, s);
    if (rc || (rc = snap.put(s))) return rc;
    m->day = day;
    m->kt = kt;
    m->ext_valid = true;
    m->int_valid = false;
    if (!m->prepared) return HYDRO_OK;
    return (m->d.dtype == HYDRO_F64) ? ensure_internal<double>(m, s) : ensure_internal<float>(m, s);
}
}  // namespace

int hydro_set_cell_coords(hydro_mesh *m, const double *xc, const double *yc) {
    if (!m || (m->CELL > 0 && (!xc || !yc))) return fail(HYDRO_ERR_INVALID, "null argument");
    for (int i = 0; i < m->CELL; i++)
        if (!std::isfinite(xc[i]) || !std::isfinite(yc[i]))
            return fail(HYDRO_ERR_INVALID, "cell %d has a non-finite coordinate", i);
    m->xc.assign(xc, xc + m->CELL);
    m->yc.assign(yc, yc + m->CELL);
    return HYDRO_OK;
}

int hydro_regroup(hydro_mesh *m, hipStream_t stream) {
    if (!m) return fail(HYDRO_ERR_INVALID, "null mesh");
    return (m->d.dtype == HYDRO_F64) ? regroup_t<double>(m, stream) : regroup_t<float>(m, stream);
}
int hydro_plan_grouping(hydro_mesh *m, int *pair_chunks, int *mixed_chunks, hipStream_t stream) {
    if (!m || !pair_chunks || !mixed_chunks) return fail(HYDRO_ERR_INVALID, "null argument");
    return (m->d.dtype == HYDRO_F64) ? plan_grouping_t<double>(m, pair_chunks, mixed_chunks, stream)
                                     : plan_grouping_t<float>(m, pair_chunks, mixed_chunks, stream);
}
int hydro_plan_stats(const hydro_mesh *m, int *tiles, int *pairs, int *singles, double *edges_per_thread) {
    if (!m) return fail(HYDRO_ERR_INVALID, "null mesh");
    if (!m->prepared) return fail(HYDRO_ERR_INVALID, "the mesh is not prepared");
    const int nt = (m->CELL + m->tile - 1) / m->tile;
    if (tiles) *tiles = nt;
    if (pairs) *pairs = m->n_pairs;
    if (singles) *singles = m->n_singles;
    if (edges_per_thread) {
        if (m->tune.scheme == HYDRO_SCHEME_TILES)
            *edges_per_thread = nt > 0 ? (double)(m->n_pairs + m->n_singles) / ((double)nt * m->tune.block) : 0.0;
        else
            *edges_per_thread = 4.0;   // the CELL scheme: every thread solves its cell's four slots
    }
    return HYDRO_OK;
}

int hydro_autotune(hydro_mesh *m, int mode, int steps, const hydro_tune *cands, int ncand, hydro_tune *best,
                   float *us_per_step, hipStream_t s) {
    if (!m || !best || steps <= 0 || ncand < 0 || (ncand > 0 && !cands)) return fail(HYDRO_ERR_INVALID, "bad arguments");
    if (mode != HYDRO_MODE_FUSED && mode != HYDRO_MODE_GRAPH && mode != HYDRO_MODE_PERSISTENT)
        return fail(HYDRO_ERR_UNSUPPORTED, "autotune times the FUSED, GRAPH or PERSISTENT mode");
    std::vector<hydro_tune> c;
    if (ncand > 0) {
        c.assign(cands, cands + ncand);
    } else {
        const hydro_tune d = default_tune();
        auto add = [&](int block, int tile, int scheme, int wps) {
            hydro_tune t = d;
            t.block = block;
            t.tile_cells = tile;
            t.scheme = scheme;
            t.waves_per_simd = wps;
            c.push_back(t);
        };
        // block x cells per tile (their ratio sets the edges per thread),
        // the 128-VGPR build, both schemes, and the cell order
        add(256, 0, HYDRO_SCHEME_TILES, 0);
        add(256, 128, HYDRO_SCHEME_TILES, 0);
        add(128, 0, HYDRO_SCHEME_TILES, 0);
        add(256, 0, HYDRO_SCHEME_TILES, 4);
        add(128, 0, HYDRO_SCHEME_CELL, 0);
        add(256, 0, HYDRO_SCHEME_CELL, 0);
        add(256, 0, HYDRO_SCHEME_TILES, 0);
        c.back().reorder = HYDRO_REORDER_RCM;
        if ((int)m->xc.size() == m->CELL) {
            add(256, 0, HYDRO_SCHEME_TILES, 0);
            c.back().reorder = HYDRO_REORDER_HILBERT;
        }
    }
    const int per = m->spd - 1;
    const int64_t remaining = per > 0 ? (int64_t)m->d.NDAYS * per - global_step(m) : 0;
    const int warm = std::min(steps, 5);
    if (remaining < steps) return fail(HYDRO_ERR_RANGE, "%d steps requested, %lld remain", steps, (long long)remaining);
    int rc = (m->d.dtype == HYDRO_F64) ? ensure_external<double>(m, s) : ensure_external<float>(m, s);
    if (rc) return rc;
    Snapshot snap;
    if ((rc = snap.take(m, s))) return rc;
    const int day0 = m->day, kt0 = m->kt;
    const hydro_tune before = m->tune;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float best_us = -1;
    int best_i = -1;
    for (size_t i = 0; i < c.size(); i++) {
        float us = -1;
        if (hydro_prepare(m, &c[i]) == HYDRO_OK && restore_snapshot(m, snap, day0, kt0, s) == HYDRO_OK &&
            hydro_step_n(m, warm, mode, s) == HYDRO_OK && restore_snapshot(m, snap, day0, kt0, s) == HYDRO_OK &&
            hipEventRecord(e0, s) == hipSuccess && hydro_step_n(m, steps, mode, s) == HYDRO_OK &&
            hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
            check_persistent(m, s) == HYDRO_OK) {   // a persistent run that left at its spin limit does not count
            float ms = 0;
            if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) us = ms * 1000.0f / (float)steps;
        }
        (void)check_persistent(m, s);   // a failed candidate's flag must not fall on the next one
        if (us_per_step) us_per_step[i] = us;
        if (us > 0 && (best_i < 0 || us < best_us)) {
            best_us = us;
            best_i = (int)i;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const hydro_tune chosen = best_i >= 0 ? c[best_i] : before;
    if ((rc = hydro_prepare(m, &chosen)) || (rc = restore_snapshot(m, snap, day0, kt0, s))) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    *best = chosen;
    if (best_i < 0) return fail(HYDRO_ERR_UNSUPPORTED, "no candidate tuning ran");
    return HYDRO_OK;
}

int hydro_reload(hydro_mesh *m, hipStream_t stream) {
    if (!m) return fail(HYDRO_ERR_INVALID, "null mesh");
    m->ext_valid = true;
    m->int_valid = false;
    if (!m->prepared) return HYDRO_OK;
    return (m->d.dtype == HYDRO_F64) ? ensure_internal<double>(m, stream) : ensure_internal<float>(m, stream);
}

int hydro_dump(hydro_mesh *m, const char *path, hipStream_t stream) {
    if (!m || !path) return fail(HYDRO_ERR_INVALID, "null argument");
    int rc = hydro_sync(m, stream);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(stream));
    const int C = m->CELL;
    const size_t rs = (size_t)m->rsize, S4 = 4 * (size_t)C;
    std::vector<char> buf(rs * S4);
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail(HYDRO_ERR_INVALID, "cannot open %s", path);
    fwrite(&C, sizeof(int32_t), 1, fp);
    const void *cellv[4] = {m->d.H, m->d.U, m->d.V, m->d.Z};
    for (const void *p : cellv) {
        if (hipMemcpy(buf.data(), p, rs * C, hipMemcpyDeviceToHost) != hipSuccess) {
            fclose(fp);
            return fail(HYDRO_ERR_HIP, "dump copy failed");
        }
        fwrite(buf.data(), rs, C, fp);
    }
    const void *slotv[7] = {m->d.SLCOS, m->d.SLSIN, m->d.SIDE, m->d.FLUX0, m->d.FLUX1, m->d.FLUX2, m->d.FLUX3};
    for (const void *p : slotv) {
        if (hipMemcpy(buf.data(), p, rs * S4, hipMemcpyDeviceToHost) != hipSuccess) {
            fclose(fp);
            return fail(HYDRO_ERR_HIP, "dump copy failed");
        }
        fwrite(buf.data(), rs, S4, fp);
    }
    fclose(fp);
    return HYDRO_OK;
}

int hydro_diagnostics(hydro_mesh *m, double *out, hipStream_t stream) {
    if (!m || !out) return fail(HYDRO_ERR_INVALID, "null argument");
    if (!is_device_ptr(out)) return fail(HYDRO_ERR_INVALID, "out is not a device pointer");
    int rc = hydro_sync(m, stream);
    if (rc) return rc;
    const size_t need = (size_t)std::max(1, diag_records(m->CELL)) * kDiagRecord * sizeof(double);
    if (m->diag_part.bytes < need && (rc = m->diag_part.alloc(need))) return rc;
    if (m->d.dtype == HYDRO_F64)
        HIP_TRY(launch_diagnostics<double>(ext_view<double>(m), m->diag_part.as<double>(), out, stream));
    else
        HIP_TRY(launch_diagnostics<float>(ext_view<float>(m), m->diag_part.as<double>(), out, stream));
    return HYDRO_OK;
}

int hydro_diagnostics_noalloc(hydro_mesh *m, double *out, hipStream_t stream) {
    if (!m || !out) return fail(HYDRO_ERR_INVALID, "null argument");
    if (!is_device_ptr(out)) return fail(HYDRO_ERR_INVALID, "out is not a device pointer");
    const int nrec = diag_records(m->CELL);
    if (nrec > kDiagScratchRecords)
        return fail(HYDRO_ERR_INVALID,
                    "hydro_diagnostics_noalloc: a mesh of %d cells needs %d partial records; its fixed buffer "
                    "holds %d (meshes of at most %d cells); use hydro_diagnostics",
                    m->CELL, nrec, kDiagScratchRecords, kDiagScratchRecords * kDiagCellsPerBlock);
    double *part = nullptr;
    HIP_TRY(diag_scratch(&part));
    int rc = hydro_sync(m, stream);
    if (rc) return rc;
    if (m->d.dtype == HYDRO_F64)
        HIP_TRY(launch_diagnostics<double>(ext_view<double>(m), part, out, stream));
    else
        HIP_TRY(launch_diagnostics<float>(ext_view<float>(m), part, out, stream));
    return HYDRO_OK;
}

int hydro_set_profiling(hydro_mesh *m, int enable) {
    if (!m) return fail(HYDRO_ERR_INVALID, "null mesh");
    m->profiling = enable != 0;
    return HYDRO_OK;
}

int hydro_last_timing(const hydro_mesh *m, double *total_ms, int *launches, int *steps) {
    if (!m) return fail(HYDRO_ERR_INVALID, "null mesh");
    if (total_ms) *total_ms = m->last_ms;
    if (launches) *launches = m->last_launches;
    if (steps) *steps = m->last_steps;
    return HYDRO_OK;
}

// ---- adjoint / tangent-linear support ---------------------------------------

// any_dtype: the misfit loops and the maps also take fp32 meshes (an fp32
// trajectory with a double adjoint); the single-step API is fp64 only.
static int adj_ready(hydro_mesh *m, bool any_dtype = false) {
    if (!m) return fail(HYDRO_ERR_INVALID, "null mesh");
    if ((m->d.dtype != HYDRO_F64 && !any_dtype) || m->d.variant != HYDRO_VARIANT_A)
        return fail(HYDRO_ERR_UNSUPPORTED, any_dtype ? "the misfit loops need a variant-A mesh"
                                                     : "adjoint and tangent-linear steps need an fp64 variant-A mesh");
    if (!m->prepared) {
        const hydro_tune t = m->tune;
        int rc = hydro_prepare(m, &t);
        if (rc) return rc;
    }
    if (m->need_zb1)
        return fail(HYDRO_ERR_UNSUPPORTED, "adjoint needs ZB1 == ZBC + HM1 (bed gradients flow through ZB1)");
    return HYDRO_OK;
}

int hydro_adj_sizes(const hydro_mesh *m, size_t *state_bytes, size_t *wsum_bytes, size_t *cellvec_bytes) {
    if (!m) return fail(HYDRO_ERR_INVALID, "null mesh");
    const size_t C = (size_t)m->CELL;
    const size_t rs = (size_t)m->rsize;
    if (state_bytes) *state_bytes = C * 4 * rs;
    if (wsum_bytes) *wsum_bytes = C * 3 * rs;
    if (cellvec_bytes) *cellvec_bytes = C * rs;
    return HYDRO_OK;
}

int hydro_state_save(hydro_mesh *m, void *state, hipStream_t s) {
    if (!m || !state) return fail(HYDRO_ERR_INVALID, "null argument");
    int rc = (m->d.dtype == HYDRO_F64) ? ensure_internal<double>(m, s) : ensure_internal<float>(m, s);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(state, m->st[m->cur].p, m->st[m->cur].bytes, hipMemcpyDeviceToDevice, s));
    return HYDRO_OK;
}

int hydro_state_load(hydro_mesh *m, const void *state, hipStream_t s) {
    if (!m || !state) return fail(HYDRO_ERR_INVALID, "null argument");
    int rc = (m->d.dtype == HYDRO_F64) ? ensure_internal<double>(m, s) : ensure_internal<float>(m, s);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(m->st[m->cur].p, state, m->st[m->cur].bytes, hipMemcpyDeviceToDevice, s));
    m->int_valid = true;
    m->ext_valid = false;
    return HYDRO_OK;
}

// first_in (optional): the first step reads its input from there (a
// checkpoint) instead of the current state buffer.
extern "C++" template <class R, int VAR>
static int record_t(hydro_mesh *m, int n, void *states, void *wsums, hipStream_t s,
                    const State<R> *first_in = nullptr) {
    int rc = ensure_internal<R>(m, s);
    if (rc) return rc;
    const size_t C = (size_t)m->CELL;
    for (int i = 0; i < n; i++) {
        FusedView<R> f = fused_view<R>(m, m->cur);
        if (i == 0 && first_in) f.Sin = first_in;
        f.rec_state = (State<R> *)states + (size_t)i * C;
        f.rec_wsum = (R *)wsums + (size_t)i * 3 * C;
        HIP_TRY((launch_fused_step<R, VAR>(f, m->kt, m->day, m->tune.block, m->tune.xcd_remap, s)));
        m->cur ^= 1;
        advance(m);
    }
    m->ext_valid = false;
    return HYDRO_OK;
}

int hydro_step_record(hydro_mesh *m, int nsteps, void *states, void *wsums, hipStream_t s) {
    if (!m || !states || !wsums || nsteps < 0) return fail(HYDRO_ERR_INVALID, "bad argument");
    const int per = m->spd - 1;
    const int64_t remaining = per > 0 ? (int64_t)m->d.NDAYS * per - global_step(m) : 0;
    if (nsteps > remaining) return fail(HYDRO_ERR_RANGE, "%d steps requested, %lld remain", nsteps, (long long)remaining);
    if (!m->prepared) {
        const hydro_tune t = m->tune;
        int rc = hydro_prepare(m, &t);
        if (rc) return rc;
    }
    if (m->d.dtype == HYDRO_F64)
        return m->d.variant ? record_t<double, 1>(m, nsteps, states, wsums, s) : record_t<double, 0>(m, nsteps, states, wsums, s);
    return m->d.variant ? record_t<float, 1>(m, nsteps, states, wsums, s) : record_t<float, 0>(m, nsteps, states, wsums, s);
}

int hydro_adjoint_step(hydro_mesh *m, int kt, int jt, const void *s_in, const void *wsum, const void *lam_out,
                       void *lam_in, void *g_fnc, void *g_zbc, hipStream_t s) {
    int rc = adj_ready(m);
    if (rc) return rc;
    if (!s_in || !wsum || !lam_out || !lam_in || !g_fnc || !g_zbc) return fail(HYDRO_ERR_INVALID, "null buffer");
    if (jt < 0 || jt >= m->d.NDAYS) return fail(HYDRO_ERR_RANGE, "jt %d", jt);
    if (!m->wbar.p && (rc = m->wbar.alloc((size_t)m->CELL * 3 * sizeof(double)))) return rc;
    if (!m->adj_flag.p) {
        if ((rc = m->adj_flag.alloc(sizeof(int)))) return rc;
        HIP_TRY(hipMemset(m->adj_flag.p, 0, sizeof(int)));
    }
    const FusedView<double> f = fused_view<double>(m, m->cur);
    HIP_TRY(launch_adjoint_step(f, m->twin.as<uint8_t>(), (const State<double> *)s_in, (const double *)wsum,
                                (const State<double> *)lam_out, (State<double> *)lam_in, m->wbar.as<double>(),
                                (double *)g_fnc, (double *)g_zbc, kt, jt, m->adj_flag.as<int>(), s));
    return HYDRO_OK;
}

int hydro_adjoint_check(hydro_mesh *m, int *unsupported) {
    if (!m || !unsupported) return fail(HYDRO_ERR_INVALID, "null argument");
    *unsupported = 0;
    if (m->adj_flag.p) HIP_TRY(hipMemcpy(unsupported, m->adj_flag.p, sizeof(int), hipMemcpyDeviceToHost));
    return HYDRO_OK;
}

int hydro_tlm_step(hydro_mesh *m, int kt, int jt, const void *s_in, const void *ds_in, const void *dfnc,
                   const void *dzbc, void *s_out, void *ds_out, hipStream_t s) {
    int rc = adj_ready(m);
    if (rc) return rc;
    if (!s_in || !ds_in || !dfnc || !dzbc || !s_out || !ds_out) return fail(HYDRO_ERR_INVALID, "null buffer");
    if (jt < 0 || jt >= m->d.NDAYS) return fail(HYDRO_ERR_RANGE, "jt %d", jt);
    const FusedView<double> f = fused_view<double>(m, m->cur);
    HIP_TRY(launch_tlm_step(f, (const State<double> *)s_in, (const State<double> *)ds_in, (const double *)dfnc,
                            (const double *)dzbc, (State<double> *)s_out, (State<double> *)ds_out, kt, jt, s));
    return HYDRO_OK;
}

static int check_gauges(hydro_mesh *m, const int32_t *cells, int n) {
    if (n < 0 || (n > 0 && !cells)) return fail(HYDRO_ERR_INVALID, "bad gauge list");
    std::vector<int32_t> h;
    int rc = d2h(h, cells, (size_t)n);
    if (rc) return rc;
    for (int i = 0; i < n; i++)
        if (h[i] < 0 || h[i] >= m->CELL) return fail(HYDRO_ERR_INVALID, "gauge cell %d out of range", h[i]);
    return HYDRO_OK;
}

int hydro_gauge_values(hydro_mesh *m, const void *state, const int32_t *cells, int n, double *out, hipStream_t s) {
    int rc = adj_ready(m, true);
    if (rc || (rc = check_gauges(m, cells, n))) return rc;
    if (m->d.dtype == HYDRO_F64)
        HIP_TRY(launch_gauge_gather((const State<double> *)state, m->inv.as<int32_t>(), cells, n, out, s));
    else
        HIP_TRY(launch_gauge_gather((const State<float> *)state, m->inv.as<int32_t>(), cells, n, out, s));
    return HYDRO_OK;
}

int hydro_gauge_seed(hydro_mesh *m, void *lam, const int32_t *cells, int n, const double *seed, hipStream_t s) {
    int rc = adj_ready(m);
    if (rc || (rc = check_gauges(m, cells, n))) return rc;
    HIP_TRY(launch_gauge_seed((State<double> *)lam, m->inv.as<int32_t>(), cells, n, seed, s));
    return HYDRO_OK;
}

// Gauge tables of the misfit loops (internal ids, per-cell first gauge and
// the chain of repeated ids), rebuilt on the stream for every call.
static int gauge_tables(hydro_mesh *m, const int32_t *gauges, int n, hipStream_t s) {
    int rc;
    if (m->gmap.bytes != (size_t)m->CELL * sizeof(int32_t) && (rc = m->gmap.alloc((size_t)m->CELL * sizeof(int32_t))))
        return rc;
    const size_t nb = (size_t)std::max(n, 1) * sizeof(int32_t);
    if (m->gnext.bytes < nb && ((rc = m->gnext.alloc(nb)) || (rc = m->gcells.alloc(nb)))) return rc;
    HIP_TRY(launch_gauge_map(m->inv.as<int32_t>(), gauges, n, m->CELL, m->gmap.as<int32_t>(), m->gnext.as<int32_t>(),
                             m->gcells.as<int32_t>(), s));
    return HYDRO_OK;
}

extern "C++" template <class R>
static int misfit_forward_t(hydro_mesh *m, int nsteps, int seg, void *ckpt, const int32_t *gauges, int n, double *zg,
                            hipStream_t s) {
    int rc;
    if ((rc = ensure_internal<R>(m, s)) || (rc = gauge_tables(m, gauges, n, s))) return rc;
    const size_t C = (size_t)m->CELL;
    const int every = regroup_every(m);
    for (int k = 0; k < nsteps; k++) {
        FusedView<R> f = fused_view<R>(m, m->cur);
        if (every > 0 && k > 0 && k % every == 0) HIP_TRY(launch_regroup<R>(f, s));   // the plan, by this state
        if (ckpt && k % seg == 0) f.rec_state = (State<R> *)ckpt + (size_t)(k / seg) * C;   // written by the step
        if (n > 0 && k > 0) {   // this step's input is the previous step's output
            f.gauge_map = m->gmap.as<int32_t>();
            f.gauge_next = m->gnext.as<int32_t>();
            f.gauge_n = n;
            f.gauge_out = zg + (size_t)(k - 1) * n;
        }
        HIP_TRY((launch_fused_step<R, 0>(f, m->kt, m->day, m->tune.block, m->tune.xcd_remap, s)));
        m->cur ^= 1;
        advance(m);
    }
    if (n > 0 && nsteps > 0)
        HIP_TRY(launch_gauge_gather(m->st[m->cur].as<State<R>>(), m->inv.as<int32_t>(), gauges, n,
                                    zg + (size_t)(nsteps - 1) * n, s));
    m->ext_valid = false;
    return HYDRO_OK;
}

int hydro_misfit_forward(hydro_mesh *m, int nsteps, int seg, void *ckpt, const int32_t *gauges, int n, double *zg,
                         hipStream_t s) {
    int rc = adj_ready(m, true);
    if (rc || (rc = check_gauges(m, gauges, n))) return rc;
    if (nsteps < 0 || seg <= 0 || (n > 0 && !zg)) return fail(HYDRO_ERR_INVALID, "bad argument");
    const int per = m->spd - 1;
    const int64_t remaining = per > 0 ? (int64_t)m->d.NDAYS * per - global_step(m) : 0;
    if (nsteps > remaining) return fail(HYDRO_ERR_RANGE, "%d steps requested, %lld remain", nsteps, (long long)remaining);
    return m->d.dtype == HYDRO_F64 ? misfit_forward_t<double>(m, nsteps, seg, ckpt, gauges, n, zg, s)
                                   : misfit_forward_t<float>(m, nsteps, seg, ckpt, gauges, n, zg, s);
}

// The adjoint steps of one recorded segment (steps s0 .. s0+len-1, records
// i = 0 .. len-1, S scenarios per record), last step first: the last step's
// cell updates are reversed from L[cur] first, then every adjoint step also
// reverses the update of the step before it.  On return L[cur] holds the
// adjoint of the segment's input state.
extern "C++" template <class R>
static int adjoint_segment(hydro_mesh *m, const FusedView<R> &f, int S, int s0, int len, int64_t g0, int n,
                           const double *resid, void *states, void *wsums, State<double> *L[2], int &cur,
                           double *g_fnc, double *g_zbc, hipStream_t s) {
    const int per = m->spd - 1;
    const size_t rec = (size_t)S * m->CELL;
    const size_t ub_rec = (size_t)kUbDoubles * rec;
    int rc;
    if (m->adj_ub.bytes < 2 * ub_rec * sizeof(double) && (rc = m->adj_ub.alloc(2 * ub_rec * sizeof(double))))
        return rc;
    double *ub[2] = {m->adj_ub.as<double>(), m->adj_ub.as<double>() + ub_rec};
    auto st = [&](int i) { return (const State<R> *)states + (size_t)i * rec; };
    auto ws = [&](int i) { return (const R *)wsums + (size_t)i * 3 * rec; };
    auto seed = [&](int i) { return n > 0 ? resid + (size_t)(s0 + i) * S * n : nullptr; };
    const int32_t *gmap = m->gmap.as<int32_t>(), *gnext = m->gnext.as<int32_t>();
    int u = 0;
    HIP_TRY(launch_adjoint_prep<R>(f, st(len - 1), ws(len - 1), seed(len - 1), L[cur], ub[u], gmap, gnext, n, S, s));
    for (int i = len - 1; i >= 0; i--) {
        const int64_t g = g0 + s0 + i;
        const bool prev = i > 0;
        HIP_TRY(launch_adjoint_step_fused<R>(f, m->twin.as<uint8_t>(), m->twin_kind.as<uint8_t>(), st(i), ws(i),
                                             prev ? st(i - 1) : nullptr,
                                             prev ? ws(i - 1) : nullptr, prev ? seed(i - 1) : nullptr, ub[u],
                                             prev ? ub[u ^ 1] : nullptr, L[cur ^ 1], g_fnc, g_zbc, gmap, gnext, n, S,
                                             1 + (int)(g % per), (int)(g / per), m->adj_flag.as<int>(), s));
        cur ^= 1;
        u ^= 1;
    }
    return HYDRO_OK;
}

// The adjoint sweep of nseg segments, last first, with the recomputation of
// segment k-1 (rec, on the mesh's side stream) overlapping the adjoint steps
// of segment k (adj, on the caller's stream): rec(k, b, stream) records
// segment k into record buffer b = k & 1, adj(k, b, stream) runs its adjoint
// steps from it.  Events order each adjoint after its segment's recording and
// each recording after the adjoint that last read its buffer; the caller's
// stream ends after all of it.  Both streams have the default priority.
extern "C++" template <class RecFn, class AdjFn>
static int pipelined_sweep(hydro_mesh *m, int nseg, hipStream_t s, RecFn rec, AdjFn adj) {
    if (nseg <= 0) return HYDRO_OK;
    if (!m->side_stream) HIP_TRY(hipStreamCreateWithFlags(&m->side_stream, hipStreamNonBlocking));
    for (hipEvent_t *e : {&m->sweep_start, &m->rec_done[0], &m->rec_done[1], &m->adj_done[0], &m->adj_done[1]})
        if (!*e) HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
    hipStream_t side = m->side_stream;
    HIP_TRY(hipEventRecord(m->sweep_start, s));   // everything the caller issued before
    HIP_TRY(hipStreamWaitEvent(side, m->sweep_start, 0));
    int rc;
    if ((rc = rec(nseg - 1, (nseg - 1) & 1, side))) return rc;
    HIP_TRY(hipEventRecord(m->rec_done[(nseg - 1) & 1], side));
    for (int k = nseg - 1; k >= 0; k--) {
        const int b = k & 1;
        if (k >= 1) {
            const int bb = (k - 1) & 1;
            if (k + 1 <= nseg - 1) HIP_TRY(hipStreamWaitEvent(side, m->adj_done[bb], 0));   // adj(k+1) read bb
            if ((rc = rec(k - 1, bb, side))) return rc;
            HIP_TRY(hipEventRecord(m->rec_done[bb], side));
        }
        HIP_TRY(hipStreamWaitEvent(s, m->rec_done[b], 0));
        if ((rc = adj(k, b, s))) return rc;
        HIP_TRY(hipEventRecord(m->adj_done[b], s));
    }
    return HYDRO_OK;
}

extern "C++" template <class R>
static int misfit_backward_t(hydro_mesh *m, int nsteps, int seg, const void *ckpt, int n, const double *resid,
                             void *states, void *wsums, void *lam, double *g_fnc, double *g_zbc, int *lam_index,
                             hipStream_t s) {
    const int per = m->spd - 1;
    const int64_t g0 = global_step(m);
    const size_t C = (size_t)m->CELL;
    State<double> *L[2] = {(State<double> *)lam, (State<double> *)lam + C};
    HIP_TRY(hipMemsetAsync(L[0], 0, C * sizeof(State<double>), s));
    int cur = 0;
    const int nseg = (nsteps + seg - 1) / seg;
    auto st_b = [&](int b) { return (void *)((State<R> *)states + (size_t)b * seg * C); };
    auto ws_b = [&](int b) { return (void *)((R *)wsums + (size_t)b * seg * 3 * C); };
    auto rec = [&](int k, int b, hipStream_t st) {
        const int s0 = k * seg, len = std::min(seg, nsteps - s0);
        // the segment again from its checkpoint (read in place by the first step)
        m->int_valid = true;
        m->ext_valid = false;
        m->day = (int)((g0 + s0) / per);
        m->kt = 1 + (int)((g0 + s0) % per);
        return record_t<R, 0>(m, len, st_b(b), ws_b(b), st, (const State<R> *)ckpt + (size_t)k * C);
    };
    auto adj = [&](int k, int b, hipStream_t st) {
        const int s0 = k * seg, len = std::min(seg, nsteps - s0);
        return adjoint_segment<R>(m, fused_view<R>(m, m->cur), 1, s0, len, g0, n, resid, st_b(b), ws_b(b), L, cur,
                                  g_fnc, g_zbc, st);
    };
    int rc = pipelined_sweep(m, nseg, s, rec, adj);
    if (rc) return rc;
    *lam_index = cur;
    return HYDRO_OK;
}

int hydro_misfit_backward(hydro_mesh *m, int nsteps, int seg, const void *ckpt, const int32_t *gauges, int n,
                          const double *resid, void *states, void *wsums, void *lam, double *g_fnc, double *g_zbc,
                          int *lam_index, hipStream_t s) {
    int rc = adj_ready(m, true);
    if (rc || (rc = check_gauges(m, gauges, n))) return rc;
    if (nsteps < 0 || seg <= 0 || !ckpt || (n > 0 && !resid) || !states || !wsums || !lam || !g_fnc || !g_zbc ||
        !lam_index)
        return fail(HYDRO_ERR_INVALID, "bad argument");
    const int per = m->spd - 1;
    if (per <= 0) return fail(HYDRO_ERR_RANGE, "the mesh has no steps per day");
    const int64_t g0 = global_step(m);
    if (nsteps > (int64_t)m->d.NDAYS * per - g0)
        return fail(HYDRO_ERR_RANGE, "%d steps requested past the NDAYS cap", nsteps);
    if ((rc = gauge_tables(m, gauges, n, s))) return rc;
    if (!m->adj_flag.p) {
        if ((rc = m->adj_flag.alloc(sizeof(int)))) return rc;
        HIP_TRY(hipMemset(m->adj_flag.p, 0, sizeof(int)));
    }
    return m->d.dtype == HYDRO_F64
               ? misfit_backward_t<double>(m, nsteps, seg, ckpt, n, resid, states, wsums, lam, g_fnc, g_zbc,
                                           lam_index, s)
               : misfit_backward_t<float>(m, nsteps, seg, ckpt, n, resid, states, wsums, lam, g_fnc, g_zbc,
                                          lam_index, s);
}

// ---- scenarios: S Manning fields on one mesh, one launch per step ---------
// The scenarios start from the mesh's current state and time, share its bed,
// geometry and gauges, and have their own FNC (fnc [S][C], caller order,
// converted to the mesh's type) and states.  The mesh's own state and time
// are left as they were.  Buffers as in the one-field loops with a scenario
// dimension inside each record: ckpt and states [record][S][C] states, wsums
// [record][S][3][C], zg and resid [step][S][n], lam [2][S][C], g_fnc and
// g_zbc [S][C] (internal order).
extern "C++" template <class R>
static int scen_setup(hydro_mesh *m, int S, const double *fnc, hipStream_t s) {
    int rc;
    const size_t C = (size_t)m->CELL;
    for (int b = 0; b < 2; b++)
        if (m->scen_st[b].bytes != (size_t)S * C * sizeof(State<R>) &&
            (rc = m->scen_st[b].alloc((size_t)S * C * sizeof(State<R>))))
            return rc;
    if (m->scen_fnc.bytes != (size_t)S * C * sizeof(R) && (rc = m->scen_fnc.alloc((size_t)S * C * sizeof(R))))
        return rc;
    HIP_TRY(launch_scen_to_internal<R>(fnc, m->orig.as<int32_t>(), m->scen_fnc.as<R>(), m->CELL, S, s));
    return HYDRO_OK;
}

extern "C++" template <class R>
static FusedView<R> scen_view(hydro_mesh *m, int cur) {
    FusedView<R> f = fused_view<R>(m, m->cur);
    f.Sin = m->scen_st[cur].as<State<R>>();
    f.Sout = m->scen_st[cur ^ 1].as<State<R>>();
    f.FNC = m->scen_fnc.as<R>();
    return f;
}

extern "C++" template <class R>
static int scen_forward_t(hydro_mesh *m, int S, const double *fnc, int nsteps, int seg, void *ckpt,
                          const int32_t *gauges, int n, double *zg, hipStream_t s) {
    int rc;
    if ((rc = ensure_internal<R>(m, s)) || (rc = scen_setup<R>(m, S, fnc, s)) ||
        (rc = gauge_tables(m, gauges, n, s)))
        return rc;
    const size_t C = (size_t)m->CELL, rec = (size_t)S * C;
    for (int k = 0; k < S; k++)
        HIP_TRY(hipMemcpyAsync(m->scen_st[0].as<State<R>>() + k * C, m->st[m->cur].p, m->st[m->cur].bytes,
                               hipMemcpyDeviceToDevice, s));
    const int day0 = m->day, kt0 = m->kt;
    const int every = regroup_every(m);
    int cur = 0;
    for (int k = 0; k < nsteps; k++) {
        FusedView<R> f = scen_view<R>(m, cur);
        // the shared plan regrouped by the first scenario's state
        if (every > 0 && k > 0 && k % every == 0) HIP_TRY(launch_regroup<R>(f, s));
        if (ckpt && k % seg == 0) f.rec_state = (State<R> *)ckpt + (size_t)(k / seg) * rec;   // written by the step
        if (n > 0 && k > 0) {
            f.gauge_map = m->gmap.as<int32_t>();
            f.gauge_next = m->gnext.as<int32_t>();
            f.gauge_n = n;
            f.gauge_out = zg + (size_t)(k - 1) * S * n;
        }
        HIP_TRY((launch_fused_step_scen<R, 0>(f, S, m->kt, m->day, m->tune.block, m->tune.xcd_remap, s)));
        cur ^= 1;
        advance(m);
    }
    if (n > 0 && nsteps > 0)
        HIP_TRY(launch_gauge_gather_scen<R>(m->scen_st[cur].as<State<R>>(), m->CELL, S, m->inv.as<int32_t>(), gauges, n,
                                            zg + (size_t)(nsteps - 1) * S * n, s));
    m->day = day0;
    m->kt = kt0;
    return HYDRO_OK;
}

extern "C++" template <class R>
static int scen_backward_t(hydro_mesh *m, int S, const double *fnc, int nsteps, int seg, const void *ckpt, int n,
                           const double *resid, void *states, void *wsums, void *lam, double *g_fnc, double *g_zbc,
                           int *lam_index, hipStream_t s) {
    int rc;
    if ((rc = ensure_internal<R>(m, s)) || (rc = scen_setup<R>(m, S, fnc, s))) return rc;
    const int per = m->spd - 1;
    const int day0 = m->day, kt0 = m->kt;
    const int64_t g0 = global_step(m);
    const size_t C = (size_t)m->CELL, rec = (size_t)S * C;
    State<double> *L[2] = {(State<double> *)lam, (State<double> *)lam + rec};
    HIP_TRY(hipMemsetAsync(L[0], 0, rec * sizeof(State<double>), s));
    int cur = 0;
    const int nseg = (nsteps + seg - 1) / seg;
    auto st_b = [&](int b) { return (State<R> *)states + (size_t)b * seg * rec; };
    auto ws_b = [&](int b) { return (R *)wsums + (size_t)b * seg * 3 * rec; };
    auto rec_fn = [&](int k, int b, hipStream_t st) -> int {
        const int s0 = k * seg, len = std::min(seg, nsteps - s0);
        m->day = (int)((g0 + s0) / per);
        m->kt = 1 + (int)((g0 + s0) % per);
        int sc = 0;
        for (int i = 0; i < len; i++) {   // the segment again, recording each step's input and flux sums
            FusedView<R> f = scen_view<R>(m, sc);
            if (i == 0) f.Sin = (const State<R> *)ckpt + (size_t)k * rec;   // the checkpoint, read in place
            f.rec_state = st_b(b) + (size_t)i * rec;
            f.rec_wsum = ws_b(b) + (size_t)i * 3 * rec;
            HIP_TRY((launch_fused_step_scen<R, 0>(f, S, m->kt, m->day, m->tune.block, m->tune.xcd_remap, st)));
            sc ^= 1;
            advance(m);
        }
        return HYDRO_OK;
    };
    auto adj_fn = [&](int k, int b, hipStream_t st) {
        const int s0 = k * seg, len = std::min(seg, nsteps - s0);
        return adjoint_segment<R>(m, scen_view<R>(m, 0), S, s0, len, g0, n, resid, st_b(b), ws_b(b), L, cur, g_fnc,
                                  g_zbc, st);
    };
    if ((rc = pipelined_sweep(m, nseg, s, rec_fn, adj_fn))) return rc;
    m->day = day0;
    m->kt = kt0;
    *lam_index = cur;
    return HYDRO_OK;
}

static int scen_ready(hydro_mesh *m, int S, const double *fnc, int nsteps, int seg) {
    int rc = adj_ready(m, true);
    if (rc) return rc;
    if (S <= 0 || !fnc || nsteps < 0 || seg <= 0) return fail(HYDRO_ERR_INVALID, "bad argument");
    if (m->tune.scheme != HYDRO_SCHEME_TILES || m->rare)
        return fail(HYDRO_ERR_UNSUPPORTED, "scenario loops need the TILES scheme and no KLAS 3/5/6/7 or unknown codes");
    const int per = m->spd - 1;
    if (per <= 0) return fail(HYDRO_ERR_RANGE, "the mesh has no steps per day");
    if (nsteps > (int64_t)m->d.NDAYS * per - global_step(m))
        return fail(HYDRO_ERR_RANGE, "%d steps requested past the NDAYS cap", nsteps);
    return HYDRO_OK;
}

int hydro_scen_misfit_forward(hydro_mesh *m, int S, const double *fnc, int nsteps, int seg, void *ckpt,
                              const int32_t *gauges, int n, double *zg, hipStream_t s) {
    int rc = scen_ready(m, S, fnc, nsteps, seg);
    if (rc || (rc = check_gauges(m, gauges, n))) return rc;
    if (n > 0 && !zg) return fail(HYDRO_ERR_INVALID, "bad argument");
    return m->d.dtype == HYDRO_F64 ? scen_forward_t<double>(m, S, fnc, nsteps, seg, ckpt, gauges, n, zg, s)
                                   : scen_forward_t<float>(m, S, fnc, nsteps, seg, ckpt, gauges, n, zg, s);
}

int hydro_scen_misfit_backward(hydro_mesh *m, int S, const double *fnc, int nsteps, int seg, const void *ckpt,
                               const int32_t *gauges, int n, const double *resid, void *states, void *wsums,
                               void *lam, double *g_fnc, double *g_zbc, int *lam_index, hipStream_t s) {
    int rc = scen_ready(m, S, fnc, nsteps, seg);
    if (rc || (rc = check_gauges(m, gauges, n))) return rc;
    if (!ckpt || (n > 0 && !resid) || !states || !wsums || !lam || !g_fnc || !g_zbc || !lam_index)
        return fail(HYDRO_ERR_INVALID, "bad argument");
    if ((rc = gauge_tables(m, gauges, n, s))) return rc;
    if (!m->adj_flag.p) {
        if ((rc = m->adj_flag.alloc(sizeof(int)))) return rc;
        HIP_TRY(hipMemset(m->adj_flag.p, 0, sizeof(int)));
    }
    return m->d.dtype == HYDRO_F64
               ? scen_backward_t<double>(m, S, fnc, nsteps, seg, ckpt, n, resid, states, wsums, lam, g_fnc, g_zbc,
                                         lam_index, s)
               : scen_backward_t<float>(m, S, fnc, nsteps, seg, ckpt, n, resid, states, wsums, lam, g_fnc, g_zbc,
                                        lam_index, s);
}

int hydro_cells_map(hydro_mesh *m, const double *in, double *out, int to_external, hipStream_t s) {
    int rc = adj_ready(m, true);
    if (rc) return rc;
    HIP_TRY(launch_cells_map(in, m->orig.as<int32_t>(), out, m->CELL, to_external, s));
    return HYDRO_OK;
}

int hydro_states_map(hydro_mesh *m, void *state, double *H, double *U, double *V, double *Z, int to_external,
                     hipStream_t s) {
    int rc = adj_ready(m, true);
    if (rc) return rc;
    HIP_TRY(launch_states_map((State<double> *)state, m->orig.as<int32_t>(), H, U, V, Z, m->CELL, to_external, s));
    return HYDRO_OK;
}

int hydro_refresh_statics(hydro_mesh *m, hipStream_t s) {
    if (!m) return fail(HYDRO_ERR_INVALID, "null mesh");
    if (!m->prepared) return HYDRO_OK;
    if (m->d.dtype == HYDRO_F64) {
        HIP_TRY(launch_refresh_statics<double>(ext_view<double>(m), m->orig.as<int32_t>(), m->ZBC.as<double>(),
                                               m->AREA.as<double>(), m->FNC.as<double>(), m->DTA.as<double>(),
                                               m->need_zb1 ? m->ZB1.as<double>() : nullptr, m->CELL, s));
    } else {
        HIP_TRY(launch_refresh_statics<float>(ext_view<float>(m), m->orig.as<int32_t>(), m->ZBC.as<float>(),
                                              m->AREA.as<float>(), m->FNC.as<float>(), m->DTA.as<float>(),
                                              m->need_zb1 ? m->ZB1.as<float>() : nullptr, m->CELL, s));
    }
    return HYDRO_OK;
}

// ---- batches ------------------------------------------------------------------

}  // extern "C"

struct hydro_batch {
    std::vector<hydro_mesh *> meshes;
    int dtype = 0, variant = 0, block = 256, remap = 1, chunk = 16;
    int regroup = 0, since_group = 0;   // hydro_tune.regroup_steps; steps since the members' plans were grouped
    unsigned spin_limit = 1u << 24;
    int ntiles = 0;
    bool rare = false;
    int scheme = 0;
    DevBuf views, tile_mesh, tile0, g0, spd, tick, barrier, perr;
    bool persist_pending = false;   // a persistent launch whose wait flag has not been read back
    std::vector<int32_t> h_tile0;
    hipStream_t cap_stream = nullptr;
    std::map<GraphKey, std::unique_ptr<GraphExec>> graphs;
    ~hydro_batch() {
        graphs.clear();
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
    }
};

namespace {

template <class R> int batch_upload(hydro_batch *b, hipStream_t s, std::vector<char> &host_views) {
    const int n = (int)b->meshes.size();
    host_views.resize(sizeof(FusedView<R>) * n);
    std::vector<int64_t> g0(n);
    std::vector<int32_t> spd(n);
    for (int i = 0; i < n; i++) {
        hydro_mesh *m = b->meshes[i];
        int rc = ensure_internal<R>(m, s);
        if (rc) return rc;
        FusedView<R> f = fused_view<R>(m, m->cur);
        std::memcpy(host_views.data() + sizeof(FusedView<R>) * i, &f, sizeof f);
        g0[i] = global_step(m);
        spd[i] = m->spd;
    }
    HIP_TRY(hipMemcpyAsync(b->views.p, host_views.data(), host_views.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(b->g0.p, g0.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(b->spd.p, spd.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));   // host_views are pageable and short-lived
    return HYDRO_OK;
}

template <class R> BatchView<R> batch_view(const hydro_batch *b) {
    BatchView<R> v;
    v.views = b->views.as<FusedView<R>>();
    v.tile_mesh = b->tile_mesh.as<int32_t>();
    v.tile0 = b->tile0.as<int32_t>();
    v.g0 = b->g0.as<int64_t>();
    v.spd = b->spd.as<int32_t>();
    v.ntiles = b->ntiles;
    v.rare = b->rare;
    v.scheme = b->scheme;
    return v;
}

// As check_persistent for one mesh: a batch persistent launch whose waits
// exceeded their spin limit left early, so the members hold partial states.
// Reported (and the flag cleared) by the next batch call that launches or
// reads the states.
int check_batch_persistent(hydro_batch *b, hipStream_t s) {
    if (!b->persist_pending) return HYDRO_OK;
    b->persist_pending = false;
    int err = 0;
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipMemcpy(&err, b->perr.p, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
        HIP_TRY(hipMemset(b->perr.p, 0, sizeof(int)));
        return fail(HYDRO_ERR_HIP, "a batch persistent launch stopped: a wait between steps exceeded its spin limit");
    }
    return HYDRO_OK;
}

void batch_advance(hydro_batch *b, int n) {
    for (hydro_mesh *m : b->meshes) {
        for (int i = 0; i < n; i++) advance(m);
        if (n & 1) m->cur ^= 1;
        m->ext_valid = false;
    }
}

template <class R, int VAR> int batch_steps(hydro_batch *b, int n, int mode, hipStream_t s, bool regroup = false) {
    std::vector<char> hv;
    int rc = check_batch_persistent(b, s);
    if (rc) return rc;
    rc = batch_upload<R>(b, s, hv);
    if (rc) return rc;
    const BatchView<R> v = batch_view<R>(b);
    if (regroup) {
        HIP_TRY(launch_batch_regroup<R>(v, s));
        b->since_group = 0;
    }
    if (mode == HYDRO_MODE_FUSED) {
        for (int i = 0; i < n; i++) HIP_TRY((launch_batch_step<R, VAR>(v, i, b->block, b->remap, s)));
    } else if (mode == HYDRO_MODE_GRAPH) {
        int chunk = std::max(2, b->chunk);
        if (chunk & 1) chunk++;
        if (!b->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
        int done = 0;
        if (n >= chunk) {
            GraphKey key{HYDRO_MODE_GRAPH, 0, chunk};
            auto it = b->graphs.find(key);
            if (it == b->graphs.end()) {
                auto ge = std::make_unique<GraphExec>();
                HIP_TRY(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
                hipError_t err = hipSuccess;
                for (int k = 0; k < chunk && err == hipSuccess; k++)
                    err = launch_batch_step_tick<R, VAR>(v, b->tick.as<int64_t>(), k, b->block, b->remap, b->cap_stream);
                if (err == hipSuccess) err = launch_tick_advance(b->tick.as<int64_t>(), chunk, b->cap_stream);
                hipError_t e2 = hipStreamEndCapture(b->cap_stream, &ge->g);
                if (err != hipSuccess || e2 != hipSuccess)
                    return fail(HYDRO_ERR_HIP, "batch graph capture failed: %s", hipGetErrorString(err ? err : e2));
                HIP_TRY(hipGraphInstantiate(&ge->x, ge->g, nullptr, nullptr, 0));
                it = b->graphs.emplace(key, std::move(ge)).first;
            }
            hipLaunchKernelGGL(k_set_tick, dim3(1), dim3(64), 0, s, b->tick.as<int64_t>(), (int64_t)0);
            HIP_TRY(hipGetLastError());
            while (n - done >= chunk) {
                HIP_TRY(hipGraphLaunch(it->second->x, s));
                done += chunk;
            }
        }
        for (int i = done; i < n; i++) HIP_TRY((launch_batch_step<R, VAR>(v, i, b->block, b->remap, s)));
    } else if (mode == HYDRO_MODE_PERSISTENT) {
        if (n > 0) {
            int maxg = 0;
            HIP_TRY((batch_persistent_max_grid<R, VAR>(b->block, b->rare, b->scheme, &maxg)));
            const int grid = std::max(1, std::min(maxg, b->ntiles));
            PersistentArgs pa;
            pa.nsteps = n;
            pa.day0 = pa.kt0 = pa.spd = 0;   // per-mesh time comes from the batch tables
            pa.flag0 = 0;
            pa.barrier = b->barrier.as<unsigned>();
            pa.error = b->perr.as<int>();
            pa.spin_limit = b->spin_limit;
            hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, pa.barrier);
            HIP_TRY(hipGetLastError());
            HIP_TRY((launch_batch_persistent<R, VAR>(v, pa, b->block, grid, s)));
            b->persist_pending = true;
        }
    } else {
        return fail(HYDRO_ERR_UNSUPPORTED, "batch mode %d (FUSED, GRAPH or PERSISTENT)", mode);
    }
    batch_advance(b, n);
    return HYDRO_OK;
}

// As steps_grouped for a batch: requests in stretches of at most
// regroup_steps (0 = kRegroupSteps, -1 = one stretch), every member's plan
// regrouped by its current state before a stretch once that many steps have
// run since the last grouping.
template <class R, int VAR> int batch_grouped(hydro_batch *b, int n, int mode, hipStream_t s) {
    int every = b->regroup == 0 ? kRegroupSteps : b->regroup;
    if (mode == HYDRO_MODE_GRAPH && every > 0) {
        const int chunk = std::max(2, b->chunk + (b->chunk & 1));
        every = std::max(chunk, every / chunk * chunk);
    }
    const bool reordered = mode == HYDRO_MODE_FUSED || mode == HYDRO_MODE_GRAPH || mode == HYDRO_MODE_PERSISTENT;
    if (every < 0 || !reordered || b->scheme != HYDRO_SCHEME_TILES || n == 0)
        return batch_steps<R, VAR>(b, n, mode, s);
    while (n > 0) {
        const int seg = std::min(n, every);
        const int rc = batch_steps<R, VAR>(b, seg, mode, s, b->since_group >= every);
        if (rc) return rc;
        b->since_group += seg;
        n -= seg;
    }
    return HYDRO_OK;
}

}  // namespace

extern "C" {

int hydro_batch_create(hydro_mesh *const *meshes, int n, const hydro_tune *tune, hydro_batch **out) {
    if (!meshes || n <= 0 || !out) return fail(HYDRO_ERR_INVALID, "bad arguments");
    *out = nullptr;
    hydro_tune t = tune ? *tune : default_tune();
    auto b = std::make_unique<hydro_batch>();
    b->dtype = meshes[0]->d.dtype;
    b->variant = meshes[0]->d.variant;
    if (t.block < 64 || t.block > kFusedMaxBlock || (t.block % 64))
        return fail(HYDRO_ERR_INVALID, "batch block %d (64..%d, multiple of 64)", t.block, kFusedMaxBlock);
    b->block = t.block;
    b->remap = t.xcd_remap;
    b->chunk = t.graph_chunk;
    b->scheme = t.scheme;
    if (t.regroup_steps < -1) return fail(HYDRO_ERR_INVALID, "regroup_steps %d (-1, 0 or more)", t.regroup_steps);
    b->regroup = t.regroup_steps;
    if (t.spin_limit > 0) b->spin_limit = (unsigned)t.spin_limit;
    if (t.tile_cells <= 0) t.tile_cells = t.block;   // members keep whole tiles: the batch balances them
    std::vector<int32_t> tile_mesh, tile0(n);
    for (int i = 0; i < n; i++) {
        hydro_mesh *m = meshes[i];
        if (!m) return fail(HYDRO_ERR_INVALID, "mesh %d is NULL", i);
        if (m->d.dtype != b->dtype || m->d.variant != b->variant)
            return fail(HYDRO_ERR_INVALID, "mesh %d: every mesh of a batch needs the same dtype and variant", i);
        for (int j = 0; j < i; j++)
            if (meshes[j] == m) return fail(HYDRO_ERR_INVALID, "mesh %d appears twice", i);
        if (!m->prepared || m->tune.block != t.block || m->tune.tile_cells != t.tile_cells ||
            m->tune.reorder != t.reorder ||
            m->tune.materialize_flux != t.materialize_flux || m->tune.scheme != t.scheme) {
            int rc = hydro_prepare(m, &t);
            if (rc) return rc;
        }
        tile0[i] = (int32_t)tile_mesh.size();
        const int nt = (m->CELL + m->tile - 1) / m->tile;
        for (int k = 0; k < nt; k++) tile_mesh.push_back(i);
        b->meshes.push_back(m);
        b->rare = b->rare || m->rare;
    }
    b->ntiles = (int)tile_mesh.size();
    b->h_tile0 = tile0;
    const size_t vsz = (b->dtype == HYDRO_F64) ? sizeof(FusedView<double>) : sizeof(FusedView<float>);
    int rc;
    if ((rc = b->views.alloc(vsz * n)) || (rc = h2d(b->tile_mesh, tile_mesh)) || (rc = h2d(b->tile0, tile0)) ||
        (rc = b->g0.alloc(sizeof(int64_t) * n)) || (rc = b->spd.alloc(sizeof(int32_t) * n)) ||
        (rc = b->tick.alloc(sizeof(int64_t))) || (rc = b->barrier.alloc(sizeof(unsigned))) ||
        (rc = b->perr.alloc(sizeof(int))))
        return rc;
    HIP_TRY(hipMemset(b->perr.p, 0, sizeof(int)));
    *out = b.release();
    return HYDRO_OK;
}

int hydro_batch_destroy(hydro_batch *b) {
    delete b;
    return HYDRO_OK;
}

int hydro_batch_step_n(hydro_batch *b, int nsteps, int mode, hipStream_t stream) {
    if (!b || nsteps < 0) return fail(HYDRO_ERR_INVALID, "bad arguments");
    for (size_t i = 0; i < b->meshes.size(); i++) {
        hydro_mesh *m = b->meshes[i];
        const int per = m->spd - 1;
        const int64_t rem = per > 0 ? (int64_t)m->d.NDAYS * per - global_step(m) : 0;
        if (nsteps > rem)
            return fail(HYDRO_ERR_RANGE, "mesh %zu: %d steps requested, %lld remain", i, nsteps, (long long)rem);
    }
    if (b->dtype == HYDRO_F64)
        return b->variant ? batch_grouped<double, 1>(b, nsteps, mode, stream)
                          : batch_grouped<double, 0>(b, nsteps, mode, stream);
    return b->variant ? batch_grouped<float, 1>(b, nsteps, mode, stream)
                      : batch_grouped<float, 0>(b, nsteps, mode, stream);
}

int hydro_batch_sync(hydro_batch *b, hipStream_t stream) {
    if (!b) return fail(HYDRO_ERR_INVALID, "null batch");
    int rc0 = check_batch_persistent(b, stream);
    if (rc0) return rc0;
    for (hydro_mesh *m : b->meshes) {
        int rc = hydro_sync(m, stream);
        if (rc) return rc;
    }
    return HYDRO_OK;
}

int hydro_precalculate(int dtype, int CELL, int NOD, const int32_t *NAP, void *XP, void *YP,
                       const void *ZBC, void *Z, const void *FNC0, double HM1, void *SIDE, void *SINF,
                       void *COSF, void *SLCOS, void *SLSIN, void *AREA, void *ZB1, void *H, void *FNC) {
    if (CELL < 0 || NOD < 0) return fail(HYDRO_ERR_INVALID, "CELL %d NOD %d", CELL, NOD);
    for (int i = 0; i < 4 * CELL; i++)
        if (NAP[i] < 0 || NAP[i] > NOD) return fail(HYDRO_ERR_INVALID, "NAP[%d] = %d outside [0, NOD]", i, NAP[i]);
    if (dtype == HYDRO_F64)
        precalculate<double>(CELL, NOD, NAP, (double *)XP, (double *)YP, (const double *)ZBC, (double *)Z,
                             (const double *)FNC0, HM1, (double *)SIDE, (double *)SINF, (double *)COSF,
                             (double *)SLCOS, (double *)SLSIN, (double *)AREA, (double *)ZB1, (double *)H,
                             (double *)FNC);
    else if (dtype == HYDRO_F32)
        precalculate<float>(CELL, NOD, NAP, (float *)XP, (float *)YP, (const float *)ZBC, (float *)Z,
                            (const float *)FNC0, (float)HM1, (float *)SIDE, (float *)SINF, (float *)COSF,
                            (float *)SLCOS, (float *)SLSIN, (float *)AREA, (float *)ZB1, (float *)H,
                            (float *)FNC);
    else
        return fail(HYDRO_ERR_INVALID, "dtype %d", dtype);
    return HYDRO_OK;
}

namespace {
int check_entries(const char *what, int n, int nfiles, const int32_t *MB, const int32_t *NN, int CELL) {
    if (n < 0 || nfiles < 0) return fail(HYDRO_ERR_INVALID, "%s: %d entries, %d files", what, n, nfiles);
    if (n > 0 && (!MB || !NN)) return fail(HYDRO_ERR_INVALID, "%s: entry arrays missing", what);
    for (int j = 0; j < n; j++)
        if (MB[j] < 1 || MB[j] > CELL)
            return fail(HYDRO_ERR_INVALID, "%s entry %d names cell %d outside 1..%d", what, j, MB[j], CELL);
    return HYDRO_OK;
}
}  // namespace

int hydro_take_boundary(const hydro_boundary_desc *d, const char *bounde_dir, void *QT, void *DQT, void *ZT,
                        void *DZT, int32_t *NHQ1, void *ZW, void *QW) {
    if (!d || !bounde_dir || !QT || !DQT || !ZT || !DZT) return fail(HYDRO_ERR_INVALID, "bad arguments");
    if (d->dtype != HYDRO_F64 && d->dtype != HYDRO_F32) return fail(HYDRO_ERR_INVALID, "dtype %d", d->dtype);
    if (d->CELL < 0 || d->NDAYS < 1 || d->rows < 1 || d->rows > d->NDAYS || d->MDT <= 0 || d->NHQ < 0)
        return fail(HYDRO_ERR_INVALID, "CELL %d NDAYS %d rows %d MDT %d NHQ %d", d->CELL, d->NDAYS, d->rows, d->MDT,
                    d->NHQ);
    const double k0 = d->dtype == HYDRO_F64 ? (double)d->MDT / d->DT : (double)((float)d->MDT / (float)d->DT);
    if (!(k0 >= 1.0)) return fail(HYDRO_ERR_INVALID, "MDT / DT = %g: no whole step per row", k0);
    int st;
    if ((st = check_entries("MBZ", d->NZ, d->NNZ0, d->MBZ, d->NNZ, d->CELL)) != HYDRO_OK) return st;
    if ((st = check_entries("MBQ", d->NQ, d->NNQ0, d->MBQ, d->NNQ, d->CELL)) != HYDRO_OK) return st;
    if ((st = check_entries("MBZQ", d->NZQ, d->NNZQ0, d->MBZQ, d->NNZQ, d->CELL)) != HYDRO_OK) return st;
    if (d->NZQ > 0 && (!d->ZBC || !d->SIDE || !d->KLAS || !NHQ1 || !ZW || !QW))
        return fail(HYDRO_ERR_INVALID, "rating curves need ZBC, SIDE, KLAS, NHQ1, ZW and QW");
    BoundaryIn in{d->CELL, d->NDAYS, d->rows, d->MDT, d->NHQ, d->DT, d->HM2, d->NZ, d->NNZ0, d->MBZ, d->NNZ,
                  d->NQ, d->NNQ0, d->MBQ, d->NNQ, d->NZQ, d->NNZQ0, d->MBZQ, d->NNZQ, d->ZBC, d->SIDE, d->KLAS};
    std::string err;
    const bool ok = d->dtype == HYDRO_F64
                        ? take_boundary<double>(in, bounde_dir, (double *)QT, (double *)DQT, (double *)ZT,
                                                (double *)DZT, NHQ1, (double *)ZW, (double *)QW, err)
                        : take_boundary<float>(in, bounde_dir, (float *)QT, (float *)DQT, (float *)ZT, (float *)DZT,
                                               NHQ1, (float *)ZW, (float *)QW, err);
    return ok ? HYDRO_OK : fail(HYDRO_ERR_INVALID, "%s", err.c_str());
}

int hydro_boundary_increments(int dtype, int CELL, int rows_v, int rows_d, int K0, const void *V, void *D) {
    if (CELL < 0 || rows_v < 0 || rows_d < 0 || K0 < 1 || (CELL > 0 && rows_d > 0 && (!V || !D)))
        return fail(HYDRO_ERR_INVALID, "CELL %d rows %d/%d K0 %d", CELL, rows_v, rows_d, K0);
    if (dtype == HYDRO_F64)
        boundary_increments<double>(CELL, rows_v, rows_d, K0, (const double *)V, (double *)D);
    else if (dtype == HYDRO_F32)
        boundary_increments<float>(CELL, rows_v, rows_d, K0, (const float *)V, (float *)D);
    else
        return fail(HYDRO_ERR_INVALID, "dtype %d", dtype);
    return HYDRO_OK;
}

int hydro_hilbert_order(int CELL, const double *xc, const double *yc, int32_t *perm) {
    if (CELL < 0 || (CELL > 0 && (!xc || !yc || !perm))) return fail(HYDRO_ERR_INVALID, "bad arguments");
    for (int i = 0; i < CELL; i++)
        if (!std::isfinite(xc[i]) || !std::isfinite(yc[i]))
            return fail(HYDRO_ERR_INVALID, "cell %d has a non-finite coordinate", i);
    std::vector<int32_t> p = hilbert_order(CELL, xc, yc);
    std::memcpy(perm, p.data(), sizeof(int32_t) * CELL);
    return HYDRO_OK;
}

int hydro_rcm_order(int CELL, const int32_t *NAC, int32_t *perm) {
    if (CELL < 0 || (!NAC && CELL) || (!perm && CELL)) return fail(HYDRO_ERR_INVALID, "bad arguments");
    for (int i = 0; i < 4 * CELL; i++)
        if (NAC[i] < 0 || NAC[i] > CELL) return fail(HYDRO_ERR_INVALID, "NAC[%d] = %d", i, NAC[i]);
    std::vector<int32_t> p = rcm_order(CELL, NAC);
    std::memcpy(perm, p.data(), sizeof(int32_t) * CELL);
    return HYDRO_OK;
}

int hydro_tile_order(int CELL, const int32_t *NAC, int block, int32_t *perm) {
    if (CELL < 0 || (!NAC && CELL) || (!perm && CELL) || block < 1) return fail(HYDRO_ERR_INVALID, "bad arguments");
    for (int i = 0; i < 4 * CELL; i++)
        if (NAC[i] < 0 || NAC[i] > CELL) return fail(HYDRO_ERR_INVALID, "NAC[%d] = %d", i, NAC[i]);
    std::vector<int32_t> p = tile_order(CELL, NAC, block, rcm_order(CELL, NAC));
    std::memcpy(perm, p.data(), sizeof(int32_t) * CELL);
    return HYDRO_OK;
}

}  // extern "C"
