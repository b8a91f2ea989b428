// kernels.hpp -- views and launch helpers shared by kernels.hip and api.cpp.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "physics.hpp"

namespace hydro {

// The caller's arrays in the reference layout (hydro_mesh_desc, typed).
template <class R> struct ExtView {
    int CELL, NDAYS, NHQ;
    Scalars<R> S;
    R *H, *U, *V, *Z, *W;
    const R *ZBC, *ZB1, *AREA, *FNC, *BF;
    const int32_t *NHQ1;
    const R *ZW, *QW, *QT, *DQT, *ZT, *DZT;
    const int32_t *NAC;
    int32_t *KLAS;
    R *KLASr;   // the caller's KLAS when it holds Real values (KLAS 7 rewrites go to both), else null
    const R *SIDE, *COSF, *SINF, *SLCOS, *SLSIN;
    R *F0, *F1, *F2, *F3;
};

template <class R> struct alignas(4 * sizeof(R)) State {
    R H, U, V, Z;
};

// The reordered layout of the fused paths.  Cell p of the new order is
// original cell orig[p].  Slot arrays are slot-major: element [j*C + p].
template <class R> struct FusedView {
    int C;
    int materialize;           // also write the caller's FLUX0..3 every step
    int scheme;                // hydro_scheme: 0 tiles (edge tasks in LDS), 1 one thread per cell
    int block;                 // workgroup size (== blockDim.x of every launch)
    int tile;                  // cells per tile (<= block; == block for the CELL scheme)
    int ntiles;                // tiles = workgroups of a one-step launch (== gridDim.x)
    int occ4;                  // TILES scheme limited to 128 VGPRs (4 waves per SIMD)
    int rare;                  // mesh has KLAS 3/5/6/7 or unknown codes: launch the full variant
    Scalars<R> S;
    // the caller's arrays, read only by boundary slots, KLAS 7 rewrites and
    // materialised fluxes; kept behind a device pointer so the common path
    // does not hold them in scalar registers
    const ExtView<R> *ext;
    // ping-pong state records {H, U, V, Z}; `in` is read, `out` written,
    // swapped every step.  One record per cell keeps a neighbour gather to
    // one 32-byte (fp64) / 16-byte (fp32) load instead of four.
    const State<R> *Sin;
    State<R> *Sout;
    const R *ZBC, *AREA, *FNC;
    const R *DTA;              // DT / AREA per cell (variant A), as the update forms it
    const R *ZB1;              // null: ZB1 == ZBC + HM1 bitwise for every cell
    const int32_t *nbr;        // [4][C] new-order neighbour, -1 none
    uint8_t *meta;             // [4][C] kind | primary bit (KLAS 7 slots are rewritten)
    const R *COSF, *SINF, *SIDE;    // [4][C]
    const R *SLCOS, *SLSIN;    // [4][C] or null when SIDE*COSF / SIDE*SINF reproduce them
    const int32_t *orig;       // [C]
    // TILES scheme: tile t (cells [t*block, (t+1)*block)) owns the task
    // entries [t*task_stride, (t+1)*task_stride), in task order:
    //   task_w   packed word (TaskWord), TASK_NONE for unused entries
    //   task_nb  new-order neighbour of a one-slot task (-1 none)
    //   task_cs  [2k] COSF, SINF of the (primary) slot
    //   task_sd  [k] SIDE of the slot (both slots of a pair have the same)
    // task_n[t]: 64-entry positions tile t uses (the rest hold TASK_NONE and
    // are not read).  The arrays hold 2*block TASK_NONE entries past the last
    // tile.
    // cell_pad[p]: bit j (0-3) set when slot j of cell p is a triangle's
    // padding slot; bit 4 (kCellHalo) when a one-slot task of another tile
    // reads cell p's state.
    int task_stride;
    const uint32_t *task_w;
    const int32_t *task_nb;
    const R *task_cs, *task_sd;
    const int32_t *task_n;
    const uint8_t *cell_pad;
    // recording (each null when off): the step's input state (for the
    // adjoint, or as a checkpoint), and its flux sums WH, WU, WV as [3][C]
    State<R> *rec_state;
    R *rec_wsum;
    // persistent TILES kernel: tiles sharing an edge with tile t are
    // tile_nb[tile_nb_off[t] .. tile_nb_off[t+1]); tile_flag[t] = steps of
    // the launch that tile t has completed
    const int32_t *tile_nb_off, *tile_nb;
    int32_t *tile_flag;        // tile t's flag at tile_flag[t * kFlagStride]
    // resident kernel: the first tile_pub[t] wavefronts of tile t hold all its
    // halo cells (kCellHalo); once they have stored them the tile's flag is
    // published
    const int32_t *tile_pub;
    // gauge read (gauge_out null when off): the thread of internal cell p
    // writes Z of the step's INPUT state to gauge_out[g] for every gauge g on
    // p (gauge_map[p] the first, -1 none; gauge_next[g] the next gauge on the
    // same cell, -1 none); scenario s writes at gauge_out + s * gauge_n
    const int32_t *gauge_map, *gauge_next;
    int gauge_n;
    double *gauge_out;
};

// Packed task word of the TILES scheme: bits 0-9 local cell a, 10-11 slot of
// a, 12-21 local cell b, 22-23 slot of b, 24 pair (both slots of an edge
// inside the tile: a primary, b mirror), 25-28 slot kind of a one-slot task
// at preparation (KLAS 7 slots are re-read: they may open), 29 halo (a
// one-slot task whose neighbour is a cell of another tile).  Tiles hold up
// to 1024 cells.
constexpr uint32_t kCellHalo = 0x10;
// Tile step flags sit 256 bytes apart: every tile polls its neighbours' flags
// every step, and flags sharing cache lines would put those polls on a few
// lines of one memory channel.
constexpr int kFlagStride = 64;
constexpr uint32_t TASK_NONE = 0xffffffffu;
constexpr uint32_t TASK_PAIR = 1u << 24;
constexpr uint32_t TASK_HALO = 1u << 29;   // one-slot task whose neighbour lies in another tile
__host__ __device__ inline uint32_t task_word(int la, int ja, int lb, int jb, bool pair, int kind) {
    return (uint32_t)la | ((uint32_t)ja << 10) | ((uint32_t)lb << 12) | ((uint32_t)jb << 22) |
           (pair ? TASK_PAIR : 0u) | ((uint32_t)kind << 25);
}
__host__ __device__ inline int task_la(uint32_t w) { return (int)(w & 0x3ff); }
__host__ __device__ inline int task_ja(uint32_t w) { return (int)((w >> 10) & 3); }
__host__ __device__ inline int task_lb(uint32_t w) { return (int)((w >> 12) & 0x3ff); }
__host__ __device__ inline int task_jb(uint32_t w) { return (int)((w >> 22) & 3); }
__host__ __device__ inline int task_kind(uint32_t w) { return (int)((w >> 25) & 0xf); }

// Time of the next step on a device clock (graph replays advance it):
// day and kt as the step loop counts them, kt in 1..spd-1.
struct StepClock {
    int day, kt;
};

constexpr int kMaxBlock = 1024;
// Workgroup size limits: the TILES-scheme kernels of one mesh run up to 1024
// threads (one tile per CU, 128 VGPRs); the CELL scheme (one thread per cell,
// more than 128 VGPRs in fp64) and the batch kernels up to 256.
constexpr int kTileMaxBlock = 1024;
constexpr int kFusedMaxBlock = 256;

// Dynamic LDS of a TILES-scheme workgroup of `block` threads: cell states,
// beds, ZB1 and the three flux-sum contributions of every slot.
inline size_t tile_lds_bytes(int block, int rsize) { return (size_t)block * 18 * (size_t)rsize; }

// Launch helpers implemented in kernels.hip (explicit instantiations there).
template <class R, int VAR>
hipError_t launch_calculate_flux(const ExtView<R> &m, int kt, int jt, int block, hipStream_t s);
template <class R, int VAR>
hipError_t launch_calculate_flux_tick(const ExtView<R> &m, const StepClock *clock, int offset, int spd, int block,
                                      hipStream_t s);
template <class R, int VAR>
hipError_t launch_update_cell(const ExtView<R> &m, int block, hipStream_t s);
template <class R, int VAR>
hipError_t launch_fused_step(const FusedView<R> &f, int kt, int jt, int block, int xcd_remap, hipStream_t s);
template <class R, int VAR>
hipError_t launch_fused_step_scen(const FusedView<R> &f, int nscen, int kt, int jt, int block, int xcd_remap,
                                  hipStream_t s);
template <class R, int VAR>
hipError_t launch_fused_step_tick(const FusedView<R> &f, const StepClock *clock, int offset, int spd, int block,
                                  int xcd_remap, hipStream_t s);
hipError_t launch_fill_i32(int32_t *p, int32_t v, int n, hipStream_t s);
hipError_t launch_tick_advance(int64_t *tick, int by, hipStream_t s);
hipError_t launch_clock_advance(StepClock *clock, int by, int spd, hipStream_t s);
hipError_t launch_clock_set(StepClock *clock, int day, int kt, hipStream_t s);

struct PersistentArgs {
    int nsteps;
    int day0, kt0, spd;
    unsigned *barrier;   // zeroed before launch
    int *error;          // set to 1 if a barrier wait exceeded its spin limit
    unsigned spin_limit;
    int flag0;           // resident kernel: every tile flag holds flag0 at launch; step s publishes flag0 + s + 1
};
// f reads f.Sin on even steps and f.Sout on odd steps.
template <class R, int VAR>
hipError_t launch_persistent(const FusedView<R> &f, const PersistentArgs &pa, int block, int grid, hipStream_t s);
template <class R, int VAR>
hipError_t persistent_max_grid(int block, bool rare, int scheme, bool occ4, int *grid);
// Resident TILES kernel: one workgroup per tile (grid = f.ntiles) for all
// pa.nsteps steps; capacity = tiles the GPU holds at once.
template <class R, int VAR>
hipError_t resident_capacity(int block, bool rare, int *tiles);
template <class R, int VAR>
hipError_t launch_resident(const FusedView<R> &f, const PersistentArgs &pa, int remap, bool plain, hipStream_t s);
// Regroup each tile's pair tasks by the wet/dry branches of the state in
// f.Sin (TILES scheme; no-op otherwise).
template <class R> hipError_t launch_regroup(const FusedView<R> &f, hipStream_t s);

// external H U V Z -> internal records (new order)
template <class R>
hipError_t launch_gather(const ExtView<R> &m, const int32_t *orig, State<R> *st, int C, hipStream_t s);
// internal records and W -> external H U V Z W
template <class R>
hipError_t launch_scatter(const ExtView<R> &m, const int32_t *orig, const State<R> *st, int C, hipStream_t s);
hipError_t launch_refresh_meta(const int32_t *KLAS_ext, const int32_t *orig, uint8_t *meta, int C,
                               hipStream_t s);
// ---- batches: many meshes advanced by one launch per step --------------------
// Each workgroup tile belongs to exactly one mesh (tile_mesh[t]); the mesh's
// tiles are [tile0[m], tile0[m] + ceil(C_m / block)).  Every mesh keeps its own
// reordered buffers; views[m] points at them (Sin = state before the batch's
// first step).  Mesh m's step s runs at global time g0[m] + s.
template <class R> struct BatchView {
    const FusedView<R> *views;
    const int32_t *tile_mesh;
    const int32_t *tile0;
    const int64_t *g0;
    const int32_t *spd;
    int ntiles;
    int rare;     // some member has KLAS 3/5/6/7 or unknown codes
    int scheme;   // hydro_scheme of every member
};
template <class R, int VAR>
hipError_t launch_batch_step(const BatchView<R> &b, int step, int block, int remap, hipStream_t s);
template <class R, int VAR>
hipError_t launch_batch_step_tick(const BatchView<R> &b, const int64_t *tick, int offset, int block, int remap,
                                  hipStream_t s);
template <class R, int VAR>
hipError_t launch_batch_persistent(const BatchView<R> &b, const PersistentArgs &pa, int block, int grid,
                                   hipStream_t s);
template <class R, int VAR>
hipError_t batch_persistent_max_grid(int block, bool rare, int scheme, int *grid);
// launch_regroup for every member of a batch (views holding current states)
template <class R> hipError_t launch_batch_regroup(const BatchView<R> &b, hipStream_t s);

// internal statics <- the caller's ZBC AREA FNC (and ZB1 when stored); DTA rebuilt from AREA
template <class R>
hipError_t launch_refresh_statics(const ExtView<R> &m, const int32_t *orig, R *ZBC, R *AREA, R *FNC, R *DTA, R *ZB1,
                                  int C, hipStream_t s);

// adjoint.hip (fp64, variant A)
hipError_t launch_adjoint_step(const FusedView<double> &f, const uint8_t *twin, const State<double> *s_in,
                               const double *wsum, const State<double> *lam_out, State<double> *lam_in, double *wbar,
                               double *g_fnc, double *g_zbc, int kt, int jt, int *unsupported, hipStream_t s);
hipError_t launch_tlm_step(const FusedView<double> &f, const State<double> *s_in, const State<double> *ds_in,
                           const double *dfnc, const double *dzbc, State<double> *s_out, State<double> *ds_out, int kt,
                           int jt, hipStream_t s);
// The fused adjoint step reads every cell's reversed update of its step from
// ub_cur (kUbDoubles doubles per cell: [8][C] per scenario) and, when
// ub_next is set, reverses the previous step's update (input s_prev, sums
// wsum_prev, gauge seeds seed_prev) into ub_next from the adjoint it forms;
// launch_adjoint_prep fills ub for a segment's last step from lam_out.
constexpr int kUbDoubles = 8;
template <class R>
hipError_t launch_adjoint_prep(const FusedView<R> &f, const State<R> *s_in, const R *wsum, const double *seed,
                               const State<double> *lam_out, double *ub, const int32_t *gmap, const int32_t *gnext,
                               int nseed, int nscen, hipStream_t s);
template <class R>
hipError_t launch_adjoint_step_fused(const FusedView<R> &f, const uint8_t *twin, const uint8_t *twin_kind,
                                     const State<R> *s_in, const R *wsum,
                                     const State<R> *s_prev, const R *wsum_prev, const double *seed_prev,
                                     const double *ub_cur, double *ub_next, State<double> *lam_in, double *g_fnc,
                                     double *g_zbc, const int32_t *gmap, const int32_t *gnext, int nseed, int nscen,
                                     int kt, int jt, int *unsupported, hipStream_t s);
template <class R>
hipError_t launch_gauge_gather_scen(const State<R> *st, int C, int nscen, const int32_t *inv, const int32_t *cells,
                                    int n, double *out, hipStream_t s);
template <class R>
hipError_t launch_scen_to_internal(const double *in, const int32_t *orig, R *out, int C, int nscen, hipStream_t s);
hipError_t launch_gauge_map(const int32_t *inv, const int32_t *cells, int n, int C, int32_t *gmap, int32_t *gnext,
                            int32_t *internal, hipStream_t s);
template <class R>
hipError_t launch_gauge_gather(const State<R> *st, const int32_t *inv, const int32_t *cells, int n, double *out,
                               hipStream_t s);
hipError_t launch_gauge_seed(State<double> *lam, const int32_t *inv, const int32_t *cells, int n, const double *seed,
                             hipStream_t s);
hipError_t launch_cells_map(const double *in, const int32_t *orig, double *out, int C, int to_external, hipStream_t s);
hipError_t launch_states_map(State<double> *st, const int32_t *orig, double *H, double *U, double *V, double *Z, int C,
                             int to_external, hipStream_t s);

}  // namespace hydro
