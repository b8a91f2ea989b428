/*
 * hydro.h -- C ABI of the MI355X shallow-water timestep library
 * (libhydro_amd.so, built from hydro-bench_amd/csrc).
 *
 * The library replaces the two-kernel timestep of hydro-cal as benchmarked by
 * chhnb/hydro-bench:
 *
 *   void CalculateFlux(const MeshView&, int kt, int jt, cudaStream_t);
 *   void UpdateCell(const MeshView&, cudaStream_t);
 *        (cuda_native_impl/hydro-cal-src/include/functors.cuh:59-61,
 *         bodies cuda_native_impl/hydro-cal-src/src/functors.cu:787-799)
 *
 * and the step loops that drive them
 *        (cuda_native_impl/benchmark.cu:87-96 runSteps,
 *         taichi_impl/F2_hydro_taichi_fp64.py:734-753 step_fn,
 *         cuda_impl/F2_driver.cu:34-44 persistent driver,
 *         cuda_native_impl/benchmark.cu:225-266 graph replay).
 *
 * Plain C types only: every array argument is a raw device pointer plus the
 * sizes carried in hydro_mesh_desc.  No function calls exit(); each returns a
 * hydro_status and hydro_last_error() gives the message (thread-local).
 *
 * Layout (the reference's external layout, kept as the drop-in contract;
 * cuda_native_impl/hydro-cal-src/src/cell.cpp:11-147, side.cpp:11-52,
 * hydro-cal/SOURCE/include/mesh/{cell,side,mesh}.hpp):
 *   cell arrays    [CELL]            H U V Z W ZBC ZB1 AREA FNC BoundaryFeature NHQ1
 *   slot arrays    [4*CELL]          slot idx = 4*cell + j, j = 0..3 (a triangle's
 *                                    4th slot is padding with SIDE = 0)
 *                                    NAC (1-based neighbour, 0 = none), KLAS,
 *                                    SIDE COSF SINF SLCOS SLSIN FLUX0..3
 *   time series    [NDAYS*CELL]      QT DQT ZT DZT, element [jt*CELL + cell]
 *   rating curves  [CELL*NHQ]        ZW QW, element [cell*NHQ + k]
 * Real arrays are double when dtype == HYDRO_F64 and float when HYDRO_F32.
 * NAC and KLAS are int32, or Real as the CUDA reference stores them
 * (hydro_mesh_desc.index_real).  Their values are small integers, so every
 * comparison the kernels make is unchanged; Real ones are converted once, at
 * create, into int32 copies the library owns, and KLAS 7 rewrites go to both.
 *
 * Threading: calls on one mesh are stream ordered and must not overlap
 * (H, U, V, Z are updated in place and KLAS 7 slots are rewritten), exactly as
 * for the reference.
 */
#ifndef HYDRO_AMD_H
#define HYDRO_AMD_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HYDRO_ABI_VERSION 5

typedef enum hydro_status {
    HYDRO_OK = 0,
    HYDRO_ERR_INVALID = 1,     /* bad argument, shape or layout */
    HYDRO_ERR_HIP = 2,         /* a HIP runtime call failed */
    HYDRO_ERR_RANGE = 3,       /* step request beyond the NDAYS x (steps_per_day-1) cap */
    HYDRO_ERR_UNSUPPORTED = 4, /* mode or option not available for this mesh */
    HYDRO_ERR_NOMEM = 5
} hydro_status;

typedef enum hydro_dtype { HYDRO_F64 = 0, HYDRO_F32 = 1 } hydro_dtype;

/* Variant A: cuda_native_impl/hydro-cal-src/src/functors.cu (default; the
 *            Taichi ports are bit-exact with it at step 1).
 * Variant B: cuda_native_impl/baselines/native_cuda_v0/fp64/hydro-cal-src/src/functors.cu
 *            (KLAS 1 uses kt, supercritical test without H<HM2, per-cell CFL
 *             DT2, +-5 m/s velocity cap). */
typedef enum hydro_variant { HYDRO_VARIANT_A = 0, HYDRO_VARIANT_B = 1 } hydro_variant;

/* Step execution modes for hydro_step_n. */
typedef enum hydro_mode {
    HYDRO_MODE_TWO_KERNEL = 0, /* calculate_flux + update_cell launches on the reference layout */
    HYDRO_MODE_FUSED = 1,      /* one fused flux+update launch per step on the reordered layout */
    HYDRO_MODE_GRAPH = 2,      /* FUSED steps captured once into a hipGraph and replayed */
    HYDRO_MODE_PERSISTENT = 3, /* one launch for all steps; a tile starts a step once the tiles
                                  it shares an edge with have finished the previous one (the
                                  CELL scheme: a grid-wide barrier between steps) */
    HYDRO_MODE_GRAPH_TWO_KERNEL = 4 /* TWO_KERNEL steps captured into a hipGraph; the flux kernel
                                       reads (day, kt) from a device clock, so replays advance time */
} hydro_mode;

/* Mesh descriptor: the reference MeshView (hydro-cal/SOURCE/include/mesh/mesh.hpp:114-142)
 * flattened into plain pointers.  All pointers are DEVICE pointers owned by the
 * caller and must stay valid until hydro_mesh_destroy.  Optional arrays may be
 * NULL when the mesh has no slot of the kind that reads them (checked at
 * create time): QT/DQT for KLAS 10, ZT/DZT for KLAS 1, ZW/QW/NHQ1 for KLAS 3,
 * BoundaryFeature for KLAS 6/7. */
typedef struct hydro_mesh_desc {
    int32_t abi_version;   /* HYDRO_ABI_VERSION */
    int32_t dtype;         /* hydro_dtype */
    int32_t variant;       /* hydro_variant */
    int32_t CELL;
    int32_t NDAYS;
    int32_t NHQ;
    int32_t MDT;           /* seconds per boundary-series row; steps_per_day = (int)(MDT / DT) */
    int32_t index_real;    /* element type of NAC and KLAS: 0 = int32 (both); HYDRO_NAC_REAL and / or
                              HYDRO_KLAS_REAL = Real (double / float by dtype), as the CUDA SideView
                              holds both (side.hpp:29-43) and the Taichi port KLAS (f64, NAC i32,
                              F2_hydro_taichi_fp64.py:46-47).  Real values must be whole numbers. */
    double DT, HM1, HM2, JL; /* converted to Real inside (float for HYDRO_F32) */

    /* state [CELL], read and written */
    void *H, *U, *V, *Z, *W;
    /* cell statics [CELL] */
    const void *ZBC, *ZB1, *AREA, *FNC, *BoundaryFeature;
    const int32_t *NHQ1;
    /* rating curves [CELL*NHQ] */
    const void *ZW, *QW;
    /* boundary series [NDAYS*CELL] */
    const void *QT, *DQT, *ZT, *DZT;
    /* slots [4*CELL] */
    const void *NAC;       /* int32, or Real with HYDRO_NAC_REAL */
    void *KLAS;            /* int32, or Real with HYDRO_KLAS_REAL; written: a KLAS 7 slot that
                              overflows becomes 0 */
    const void *SIDE, *COSF, *SINF, *SLCOS, *SLSIN;
    void *FLUX0, *FLUX1, *FLUX2, *FLUX3; /* written by calculate_flux */
} hydro_mesh_desc;

#define HYDRO_NAC_REAL 1
#define HYDRO_KLAS_REAL 2

typedef struct hydro_mesh hydro_mesh; /* opaque handle */

/* Tuning of the reordered (FUSED / GRAPH / PERSISTENT) path. */
/* Cell order of the internal layout (tiles are runs of consecutive cells in
 * it).  NONE: the caller's order.  RCM: reverse Cuthill-McKee of the cell
 * graph.  HILBERT: cell centroids along a Hilbert space-filling curve (needs
 * hydro_set_cell_coords).  TILES: compact breadth-first tiles of `block`
 * cells seeded along the RCM order, so most edges of a tile join two of its
 * own cells (the TILES scheme solves those edges once).  Every order gives
 * the same results bit for bit: the primary side of an edge is decided on the
 * caller's cell ids whatever the order. */
typedef enum hydro_reorder {
    HYDRO_REORDER_NONE = 0,
    HYDRO_REORDER_RCM = 1,
    HYDRO_REORDER_HILBERT = 2,
    HYDRO_REORDER_TILES = 3
} hydro_reorder;

/* Work split of the FUSED / GRAPH / PERSISTENT step kernels.
 *  TILES: a workgroup owns a tile of `block` cells; it stages their states
 *         in LDS, solves each interior edge inside the tile once for both
 *         slots (other slots one by one), keeps the slot contributions in
 *         LDS, then each thread sums its cell's four slots in slot order and
 *         updates the cell.
 *  CELL:  one thread per cell computes its four slot fluxes (each interior
 *         edge is solved from both sides) and updates the cell.
 * Both give the same values bit for bit. */
typedef enum hydro_scheme { HYDRO_SCHEME_TILES = 0, HYDRO_SCHEME_CELL = 1 } hydro_scheme;

#define HYDRO_TILE_BALANCED (-1)

typedef struct hydro_tune {
    int32_t block;          /* threads per workgroup of the fused step (multiple of 64; 64..1024 for
                               the TILES scheme, 64..256 for the CELL scheme and for batches) */
    int32_t reorder;        /* hydro_reorder */
    int32_t xcd_remap;      /* 1: map workgroups to XCDs in contiguous tile ranges */
    int32_t graph_chunk;    /* steps per captured hipGraph (GRAPH modes) */
    int32_t persistent_wgs_per_cu; /* 0 = occupancy maximum */
    int32_t materialize_flux; /* 1: FUSED modes also write FLUX0..3 every step */
    int32_t waves_per_simd; /* 0: compiler's choice; 4: fused kernels limited to 128 VGPRs so
                               4 waves fit per SIMD (one resident round at 207K cells) */
    int32_t scheme;         /* hydro_scheme */
    int32_t tile_cells;     /* TILES scheme: cells per tile (64..block); 0 = block, or for a mesh
                               with fewer tiles than CUs one tile per CU (at least 64 cells);
                               HYDRO_TILE_BALANCED (-1) = as many tiles as the GPU holds
                               workgroups of `block` threads at once (1024/block per CU), so
                               every CU gets the same number; batches use block.  With `block`
                               threads per tile this sets the edges per thread
                               (hydro_plan_stats) */
    int32_t spin_limit;     /* PERSISTENT modes: polls (grid barrier: sleeps) a wait between steps
                               may take before the launch gives up and reports HYDRO_ERR_HIP
                               (0 = 2^24) */
    int32_t persistent_loop; /* PERSISTENT mode, TILES scheme: 0 = one resident workgroup per tile
                               for the whole launch (tile statics and task entries held on the CU
                               across steps) whenever every tile fits on the GPU at once, else the
                               loop below; 1 = workgroups loop over tiles every step */
    int32_t resident_launch; /* the resident kernel's launch: 0 = cooperative (the runtime
                               guarantees every workgroup is resident); 1 = plain launch, about
                               10 us sooner to start: the library has checked that every tile
                               fits on the GPU at once, and the caller guarantees that no kernel
                               of another stream or process holds CUs while it runs (two resident
                               launches sharing the GPU could each wait for the other's CUs until
                               the spin limit) */
    int32_t regroup_steps;  /* TILES scheme, FUSED / GRAPH / PERSISTENT modes (single meshes and
                               batches): a request is run in launches (PERSISTENT) or stretches of at
                               most this many steps (GRAPH: whole replays), and before one the plan's
                               tasks are regrouped by the current wet/dry state (hydro_regroup; every
                               member of a batch) once this many steps have run since the last
                               grouping; 0 = 200, -1 = never (one launch per request) */
} hydro_tune;

/* ---- lifetime ------------------------------------------------------------ */

/* Validate the descriptor and build the handle.  Does not launch work; the
 * caller's FLUX0..3 must already hold zeros, as cuda_native_impl side.cpp:208-211
 * and F2_hydro_taichi_fp64.py:768-771 guarantee for the reference. */
int hydro_mesh_create(const hydro_mesh_desc *desc, hydro_mesh **out);
int hydro_mesh_destroy(hydro_mesh *m);

/* Build the reordered internal layout used by the FUSED, GRAPH and
 * PERSISTENT modes (host-side reorder + device upload, synchronous).
 * tune may be NULL for defaults.  The task plan groups each tile's interior
 * edges by the wet/dry branch they take in the current state (H, Z); the
 * results do not depend on it, only the speed: calling it again once the
 * wet area has moved regroups them. */
int hydro_prepare(hydro_mesh *m, const hydro_tune *tune);

/* Autotuner (no counterpart in the reference): times each candidate tuning of
 * the reordered path on this mesh (prepare, up to 5 warm-up steps, then
 * `steps` timed steps in `mode` = FUSED, GRAPH or PERSISTENT, every candidate
 * from the same state and time), then leaves the mesh prepared with the
 * fastest, its state and time as they were, and writes that tuning to *best.
 * cands == NULL with ncand 0 tries a built-in list over block (128, 256),
 * cells per tile (edges per thread), the 128-VGPR build, both schemes and the
 * cell order (TILES, RCM, and HILBERT when coordinates are set); us_per_step
 * (optional, one entry per candidate) receives each time, -1 for a
 * candidate that could not run.  Needs `steps` steps left before the NDAYS
 * cap; synchronises the stream. */
int hydro_autotune(hydro_mesh *m, int mode, int steps, const hydro_tune *cands, int ncand, hydro_tune *best,
                   float *us_per_step, hipStream_t stream);

/* Regroup the tasks of the prepared plan by the wet/dry branches their edges
 * take in the current state, on the device (stream-ordered, one small launch;
 * no counterpart in the reference).  hydro_prepare groups them by the state
 * it finds; as the wet area moves the chunks mix again and the reordered
 * modes slow down (~6% over 900 steps of the 207K floodplain), so a long run
 * calls this every few hundred steps.  Each tile's pair tasks are permuted
 * among the entries that hold pair tasks; results do not change.  No-op for
 * the CELL scheme. */
int hydro_regroup(hydro_mesh *m, hipStream_t stream);

/* Count of the plan's pair chunks (64-entry chunks of pair tasks) and of
 * those that mix pairs of different wet/dry classes in the current state
 * (host-side, synchronises the stream; a diagnostic of hydro_regroup). */
int hydro_plan_grouping(hydro_mesh *m, int *pair_chunks, int *mixed_chunks, hipStream_t stream);

/* Cell centroids (HOST arrays [CELL], caller order) for the HILBERT order;
 * no counterpart in the reference.  Copied; the next hydro_prepare uses them. */
int hydro_set_cell_coords(hydro_mesh *m, const double *xc, const double *yc);

/* The prepared plan of the TILES scheme: tiles (= workgroups of a step),
 * edge tasks solved for both slots at once (pairs) and one-slot tasks, and
 * edges per thread ((pairs + singles) / (tiles * block); 4 slots per thread
 * for the CELL scheme).  The autotuner's cells-per-tile axis moves it. */
int hydro_plan_stats(const hydro_mesh *m, int *tiles, int *pairs, int *singles, double *edges_per_thread);

/* ---- the reference operator API ----------------------------------------- */

/* One flux pass: replaces CalculateFlux(mesh, kt, jt, stream)
 * (cuda_native_impl/hydro-cal-src/include/functors.cuh:59, functors.cu:787-792). */
int hydro_calculate_flux(hydro_mesh *m, int kt, int jt, hipStream_t stream);

/* One cell update: replaces UpdateCell(mesh, stream)
 * (functors.cuh:61, functors.cu:794-799). */
int hydro_update_cell(hydro_mesh *m, hipStream_t stream);

/* ---- step loop ----------------------------------------------------------- */

/* Run nsteps steps following the reference's (day, kt) order
 * (benchmark.cu:87-96: for day < NDAYS, for kt = 1 .. steps_per_day-1)
 * starting from the mesh's time cursor, which then advances.  Returns
 * HYDRO_ERR_RANGE without running anything if fewer than nsteps steps remain
 * (the reference's NDAYS cap, scripts/check_correctness.py:197-225). */
int hydro_step_n(hydro_mesh *m, int nsteps, int mode, hipStream_t stream);

int hydro_set_time(hydro_mesh *m, int day, int kt);
int hydro_get_time(const hydro_mesh *m, int *day, int *kt, int *steps_per_day);

/* After FUSED/GRAPH/PERSISTENT steps the current state lives in the reordered
 * internal buffers.  hydro_sync writes it back to the caller's H U V Z W (and
 * KLAS for rewritten KLAS 7 slots, and FLUX0..3 when materialize_flux is set).
 * hydro_reload goes the other way, after the caller rewrote H U V Z in place.
 * Both are stream ordered; switching modes calls them implicitly. */
int hydro_sync(hydro_mesh *m, hipStream_t stream);
int hydro_reload(hydro_mesh *m, hipStream_t stream);

/* ---- diagnostics ---------------------------------------------------------- */

const char *hydro_last_error(void);
int hydro_abi_version(void);

/* Dump in the reference benchmark's binary format (benchmark.cu:142-161):
 * int32 CELL; Real H,U,V,Z [CELL]; Real SLCOS,SLSIN,SIDE [4*CELL];
 * Real FLUX0..3 [4*CELL].  Synchronizes the stream. */
int hydro_dump(hydro_mesh *m, const char *path, hipStream_t stream);

/* Per-launch timing of the dominant kernel over the last hydro_step_n call,
 * measured with hipEvents on the launch stream (FUSED/GRAPH/PERSISTENT).
 * Enabled by hydro_set_profiling(m, 1); adds one event pair per launch. */
int hydro_set_profiling(hydro_mesh *m, int enable);
int hydro_last_timing(const hydro_mesh *m, double *total_ms, int *launches, int *steps);

/* ---- diagnostics ---------------------------------------------------------- */
/* Conservation and health of the current state, reduced on the device with
 * wavefront (DPP) reductions: the quantities of the reference's alignment
 * checker (scripts/check_correctness.py:620-655 _conservation_metrics and
 * :980-990), there computed with Kahan sums on the host.  out is a DEVICE
 * array of HYDRO_DIAG_N doubles:
 *   [HYDRO_DIAG_MASS]        sum H * AREA
 *   [HYDRO_DIAG_MOM_X/Y]     sum H * U * AREA, sum H * V * AREA
 *   [HYDRO_DIAG_KINETIC]     sum 0.5 * H * (U^2 + V^2) * AREA
 *   [HYDRO_DIAG_POTENTIAL]   sum 0.5 * 9.81 * H^2 * AREA
 *   [HYDRO_DIAG_KLAS10_INFLOW] sum over KLAS 10 slots of FLUX0 * SIDE
 *   [HYDRO_DIAG_KLAS1_INFLOW]  sum over KLAS 1 slots of FLUX0 * SIDE
 *                            (outward positive, of the caller's FLUX0: the
 *                            last flux pass of a two-kernel step, or of a
 *                            reordered step with materialize_flux)
 *   [HYDRO_DIAG_NONFINITE_H] cells whose H is NaN or infinite
 *   [HYDRO_DIAG_INF_H]       cells whose H is infinite
 *   [HYDRO_DIAG_H_MIN/MAX]   range of the finite depths (NaN if none)
 * Terms are formed in double from the stored values in the checker's order of
 * operations and summed in double-double, so each sum is the correctly
 * rounded sum of its terms (up to near-total cancellation), independent of
 * the workgroup count.  Writes the state back to the caller's arrays first
 * (as hydro_sync).  The first call allocates a partial-sum buffer of 144
 * bytes per 1024 cells. */
enum {
    HYDRO_DIAG_MASS = 0,
    HYDRO_DIAG_MOM_X = 1,
    HYDRO_DIAG_MOM_Y = 2,
    HYDRO_DIAG_KINETIC = 3,
    HYDRO_DIAG_POTENTIAL = 4,
    HYDRO_DIAG_KLAS10_INFLOW = 5,
    HYDRO_DIAG_KLAS1_INFLOW = 6,
    HYDRO_DIAG_NONFINITE_H = 7,
    HYDRO_DIAG_INF_H = 8,
    HYDRO_DIAG_H_MIN = 9,
    HYDRO_DIAG_H_MAX = 10,
    HYDRO_DIAG_N = 11
};
int hydro_diagnostics(hydro_mesh *m, double *out, hipStream_t stream);

/* hydro_diagnostics without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels and results; the partial sums go to a fixed buffer compiled into
 * the library (144 KiB per GPU, held until the process ends), so meshes of
 * more than 1,048,576 cells are rejected with HYDRO_ERR_INVALID.  All calls
 * on one GPU share that buffer: calls that overlap on one GPU (other streams,
 * other threads) race and must be ordered by the caller.  Every record a call
 * reads it has written first, so nothing needs clearing between calls. */
int hydro_diagnostics_noalloc(hydro_mesh *m, double *out, hipStream_t stream);

/* ---- batches of meshes ---------------------------------------------------- */
/* No counterpart in the reference (it runs one mesh per process): many
 * meshes of uneven size, or many scenarios as separate meshes, advanced in
 * lockstep by one launch per step (FUSED), a replayed hipGraph (GRAPH) or a
 * single persistent launch (PERSISTENT).  Workgroup tiles are packed CSR-style
 * (each tile belongs to one mesh), so small meshes share the chip instead of
 * each launching a mostly idle grid.  Every mesh keeps its own buffers and
 * time cursor; all meshes need the same dtype and variant, and are
 * (re)prepared with `tune`.  A mesh must not be stepped on its own while it
 * belongs to a live batch on another stream. */
typedef struct hydro_batch hydro_batch;
int hydro_batch_create(hydro_mesh *const *meshes, int n, const hydro_tune *tune, hydro_batch **out);
int hydro_batch_destroy(hydro_batch *b);
int hydro_batch_step_n(hydro_batch *b, int nsteps, int mode, hipStream_t stream);
int hydro_batch_sync(hydro_batch *b, hipStream_t stream);

/* ---- adjoint and tangent-linear steps (variant A) -------------------------- */
/* No counterpart in the reference: these make gradients of a gauge misfit
 * with respect to per-cell Manning n (through FNC = 9.81 n^2) and bed ZBC
 * available.  They work on the reordered layout; every buffer is caller-owned
 * device memory of the size hydro_adj_sizes reports, in the library's
 * internal cell order:
 *   state   {H, U, V, Z} per cell       (state_bytes)
 *   wsum    one step's flux sums [3][C] (wsum_bytes)
 *   cellvec one value per cell          (cellvec_bytes)
 * A step's time (kt, jt) is passed explicitly (the step that produced the
 * state after it).  The single-step calls need an fp64 mesh; the misfit
 * loops, the gauge reads and the maps also take an fp32 mesh, whose state and
 * wsum buffers are then fp32 (the trajectory) while adjoint buffers, gauge
 * values, residuals and gradients stay double. */
int hydro_adj_sizes(const hydro_mesh *m, size_t *state_bytes, size_t *wsum_bytes, size_t *cellvec_bytes);
/* current state <-> a state buffer (checkpoints) */
int hydro_state_save(hydro_mesh *m, void *state, hipStream_t stream);
int hydro_state_load(hydro_mesh *m, const void *state, hipStream_t stream);
/* nsteps fused steps from the current state and time; states + k*state_bytes
 * receives the input state of step k, wsums + k*wsum_bytes its flux sums */
int hydro_step_record(hydro_mesh *m, int nsteps, void *states, void *wsums, hipStream_t stream);
/* lam_in = (d state_out / d state_in)^T lam_out for the step (kt, jt) whose
 * input state and flux sums were recorded; g_fnc, g_zbc += parameter parts.
 * Slot kinds without an adjoint (KLAS 3, 5, 6, 7, unknown codes) raise a
 * flag read by hydro_adjoint_check. */
int hydro_adjoint_step(hydro_mesh *m, int kt, int jt, const void *s_in, const void *wsum, const void *lam_out,
                       void *lam_in, void *g_fnc, void *g_zbc, hipStream_t stream);
int hydro_adjoint_check(hydro_mesh *m, int *unsupported);
/* tangent-linear step: s_out = step(s_in), ds_out = d step . (ds_in, dfnc, dzbc) */
int hydro_tlm_step(hydro_mesh *m, int kt, int jt, const void *s_in, const void *ds_in, const void *dfnc,
                   const void *dzbc, void *s_out, void *ds_out, hipStream_t stream);
/* gauges are ORIGINAL cell ids (device int32[n]): read Z of a state buffer,
 * or add seed[i] to the Z component of an adjoint buffer */
int hydro_gauge_values(hydro_mesh *m, const void *state, const int32_t *cells, int n, double *out,
                       hipStream_t stream);
int hydro_gauge_seed(hydro_mesh *m, void *lam, const int32_t *cells, int n, const double *seed, hipStream_t stream);
/* per-cell vectors and state buffers between internal and caller order */
int hydro_cells_map(hydro_mesh *m, const double *in, double *out, int to_external, hipStream_t stream);
int hydro_states_map(hydro_mesh *m, void *state, double *H, double *U, double *V, double *Z, int to_external,
                     hipStream_t stream);
/* re-read ZBC, AREA, FNC (and ZB1 when stored) from the caller's arrays after
 * the caller changed them (e.g. a calibration update); DT / AREA of the fused
 * steps is rebuilt from the new AREA, so every mode and the adjoint agree */
int hydro_refresh_statics(hydro_mesh *m, hipStream_t stream);
/* Whole-run loops of a gauge misfit, issued from the library with no host
 * synchronisation inside (the per-call helpers above check the gauge list on
 * the host each time).
 * hydro_misfit_forward: nsteps fused steps from the current state and time;
 *   before step k with k % seg == 0 the state goes to
 *   ckpt + (k / seg) * state_bytes (ckpt may be NULL); after step k the level
 *   Z of the n gauges (original ids) goes to zg[k * n + i].
 * hydro_misfit_backward: the adjoint sweep of such a run.  The mesh's time
 *   must be the forward run's start time.  resid[k * n + i] = dL/dZ of gauge
 *   i after step k; states and wsums hold 2 * seg records each
 *   (hydro_adj_sizes): the recomputation of one segment (on a stream of the
 *   library's, default priority) overlaps the adjoint steps of the next;
 *   lam holds two state buffers, the first of which is cleared here; g_fnc
 *   and g_zbc are accumulated (internal order).  On return *lam_index (0 or
 *   1) names the buffer holding the adjoint of the initial state. */
int hydro_misfit_forward(hydro_mesh *m, int nsteps, int seg, void *ckpt, const int32_t *gauges, int n, double *zg,
                         hipStream_t stream);
int hydro_misfit_backward(hydro_mesh *m, int nsteps, int seg, const void *ckpt, const int32_t *gauges, int n,
                          const double *resid, void *states, void *wsums, void *lam, double *g_fnc, double *g_zbc,
                          int *lam_index, hipStream_t stream);
/* The same loops for S scenarios of the mesh at once (S Manning fields; the
 * bed, geometry, gauges, start state and time are shared): every step is one
 * launch over all scenarios' tiles, every adjoint step one launch over all
 * scenarios' cells.  fnc [S][CELL] is FNC = 9.81 n^2 in the caller's cell
 * order (double; converted to the mesh's type).  The mesh's own state and
 * time are left unchanged.  Each record of ckpt and states holds S states
 * [S][state_bytes], of wsums S flux sums; zg and resid are [step][S][n],
 * lam [2][S] adjoint states, g_fnc and g_zbc [S][CELL] (internal order).
 * Needs the TILES scheme and a mesh without KLAS 3/5/6/7 or unknown codes.
 * No counterpart in the reference (batched calibration hypotheses). */
int hydro_scen_misfit_forward(hydro_mesh *m, int S, const double *fnc, int nsteps, int seg, void *ckpt,
                              const int32_t *gauges, int n, double *zg, hipStream_t stream);
int hydro_scen_misfit_backward(hydro_mesh *m, int S, const double *fnc, int nsteps, int seg, const void *ckpt,
                               const int32_t *gauges, int n, const double *resid, void *states, void *wsums,
                               void *lam, double *g_fnc, double *g_zbc, int *lam_index, hipStream_t stream);

/* ---- host-side mesh preparation ------------------------------------------ */

/* Geometry preprocessing of MeshData::preCalculate
 * (cuda_native_impl/hydro-cal-src/src/mesh.cpp:280-372), on HOST arrays, in
 * Real = double (dtype HYDRO_F64) or float (HYDRO_F32; the fp32 reference
 * runs this step in float too, taichi_impl/mesh_loader_f1.py:38-44).
 *   NAP [4*CELL] 1-based node ids per cell (4th = 0 or = 1st for a triangle)
 *   XP, YP [NOD] node coordinates, shifted in place by their minima
 *   ZBC, FNC0 [CELL] bed and Manning n; Z [CELL] initial level, rewritten
 *   for dry cells
 * Outputs SIDE SINF COSF SLCOS SLSIN [4*CELL], AREA ZB1 H FNC [CELL]. */
int hydro_precalculate(int dtype, int CELL, int NOD, const int32_t *NAP, void *XP, void *YP,
                       const void *ZBC, void *Z, const void *FNC0, double HM1, void *SIDE, void *SINF,
                       void *COSF, void *SLCOS, void *SLSIN, void *AREA, void *ZB1, void *H, void *FNC);

/* Inputs of hydro_take_boundary: the boundary entries of MBZ.DAT, MBQ.DAT and
 * MBZQ.DAT as MeshData::loadFromFileMBZ/MBQ/MBZQ hold them
 * (cuda_native_impl/hydro-cal-src/src/mesh.cpp:134-217) -- per entry j a
 * 1-based cell id MB*[j] and the number NN*[j] of its series file -- and the
 * scalars the series need.  Host arrays. */
typedef struct hydro_boundary_desc {
    int32_t dtype;      /* HYDRO_F64 or HYDRO_F32: the Real of every output */
    int32_t CELL;
    int32_t NDAYS;      /* rows of the series (TIME.DAT) */
    int32_t rows;       /* rows written, 1..NDAYS; the increments of the last
                           one still come from the row after it, if any */
    int32_t MDT;        /* seconds per row */
    int32_t NHQ;        /* points per rating curve: row length of ZW / QW */
    double DT, HM2;     /* step (K0 = (int)(MDT / DT) steps per row), minimum depth */
    int32_t NZ, NNZ0;   /* level entries, level series files NZ/NZ0001.DAT .. */
    const int32_t *MBZ, *NNZ;
    int32_t NQ, NNQ0;   /* discharge entries, files NQ/NQ0001.DAT .. */
    const int32_t *MBQ, *NNQ;
    int32_t NZQ, NNZQ0; /* rating-curve entries, files NZQ/NZQ0001.DAT .. */
    const int32_t *MBZQ, *NNZQ;
    const void *ZBC;    /* [CELL] bed level (rating curves only) */
    const void *SIDE;   /* [4*CELL] side lengths, as hydro_precalculate gives them */
    const int32_t *KLAS;/* [4*CELL] slot kinds (rating curves only) */
} hydro_boundary_desc;

/* MeshData::take_boundary_for_two_d and BOUNDRYinterp
 * (cuda_native_impl/hydro-cal-src/src/mesh.cpp:374-533) with the expansion to
 * cells of CellView's constructor (cell.cpp:92-126), on HOST arrays: reads
 * <bounde_dir>/NZ/NZnnnn.DAT, NQ/NQnnnn.DAT and NZQ/NZQnnnn.DAT (a count, then
 * `time value` or `level discharge` lines, parsed into Real) and writes
 *   ZT, DZT, QT, DQT [rows*CELL] column-major: row jt of cell c at jt*CELL+c.
 *     Row jt is the series at jt * MDT seconds, linearly interpolated and
 *     rounded through float; a discharge is split evenly, in float, over the
 *     entries of its file; D = (row jt+1 - row jt) / K0, 0 on row NDAYS-1.
 *     Cells with no entry hold 0; of two entries of one cell the later wins.
 *   NHQ1 [CELL], ZW, QW [CELL*NHQ] (may be NULL when NZQ == 0): per entry cell
 *     the number of points of its curve, the levels, and its share of each
 *     point's discharge by wetted width (depth above ZBC, at least HM2, times
 *     the side of its last KLAS 3 slot).
 * Every output is overwritten in full.  A missing or malformed file, an entry
 * naming a cell outside 1..CELL or a curve longer than NHQ is
 * HYDRO_ERR_INVALID (the reference exits instead).  Bracket search differs in
 * one place: the reference also tests the bracket that ends one entry past a
 * series' table, reading outside it; that bracket is not tested here. */
int hydro_take_boundary(const hydro_boundary_desc *d, const char *bounde_dir, void *QT, void *DQT, void *ZT,
                        void *DZT, int32_t *NHQ1, void *ZW, void *QW);

/* Increments of an expanded series (host arrays, Real per dtype):
 * D[jt*CELL+c] = (V[(jt+1)*CELL+c] - V[jt*CELL+c]) / K0 for jt < rows_d, and
 * 0 where V has no row jt+1 (mesh.cpp:449-464).  For callers that build the
 * series rows themselves. */
int hydro_boundary_increments(int dtype, int CELL, int rows_v, int rows_d, int K0, const void *V, void *D);

/* Hilbert-curve order of cell centroids (host arrays): perm[new] = old.  The
 * HILBERT layout of hydro_prepare uses it. */
int hydro_hilbert_order(int CELL, const double *xc, const double *yc, int32_t *perm);

/* Reverse Cuthill-McKee order of the cell graph of a host NAC array:
 * perm[new] = old.  The reordered layout of hydro_prepare uses it. */
int hydro_rcm_order(int CELL, const int32_t *NAC, int32_t *perm);

/* The TILES cell order of hydro_prepare for tiles of `block` cells: compact
 * breadth-first tiles seeded along the RCM order, every tile but the last
 * full.  perm[new] = old. */
int hydro_tile_order(int CELL, const int32_t *NAC, int block, int32_t *perm);

#ifdef __cplusplus
}
#endif
#endif /* HYDRO_AMD_H */
