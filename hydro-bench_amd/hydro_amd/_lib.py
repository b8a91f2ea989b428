"""ctypes binding of libhydro_amd.so (the C ABI in include/hydro.h).

The shared library is built in-tree by ``make -C hydro-bench_amd`` (or
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there
is no fallback implementation of the timestep.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhydro_amd.so")

ABI_VERSION = 5

# hydro_status
OK, ERR_INVALID, ERR_HIP, ERR_RANGE, ERR_UNSUPPORTED, ERR_NOMEM = range(6)
# hydro_dtype
F64, F32 = 0, 1
# hydro_variant
VARIANT_A, VARIANT_B = 0, 1
# hydro_mesh_desc.index_real bits
NAC_REAL, KLAS_REAL = 1, 2
# hydro_mode
MODE_TWO_KERNEL, MODE_FUSED, MODE_GRAPH, MODE_PERSISTENT, MODE_GRAPH_TWO_KERNEL = range(5)
MODES = {
    "two_kernel": MODE_TWO_KERNEL,
    "fused": MODE_FUSED,
    "graph": MODE_GRAPH,
    "persistent": MODE_PERSISTENT,
    "graph_two_kernel": MODE_GRAPH_TWO_KERNEL,
}
# hydro_diagnostics record (include/hydro.h HYDRO_DIAG_*)
DIAG_FIELDS = ("mass", "momentum_x", "momentum_y", "kinetic_energy", "potential_energy", "klas10_inflow",
               "klas1_inflow", "nonfinite_h", "inf_h", "h_min", "h_max")
# hydro_reorder
REORDER_NONE, REORDER_RCM, REORDER_HILBERT, REORDER_TILES = range(4)
SCHEME_TILES, SCHEME_CELL = range(2)
TILE_BALANCED = -1   # hydro_tune.tile_cells: as many tiles as the GPU holds workgroups at once


class HydroError(RuntimeError):
    """A hydro_* call returned a non-zero hydro_status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"hydro status {status}: {message}")
        self.status = status


class MeshDesc(ctypes.Structure):
    """hydro_mesh_desc (include/hydro.h)."""

    _fields_ = [
        ("abi_version", c_int32), ("dtype", c_int32), ("variant", c_int32), ("CELL", c_int32),
        ("NDAYS", c_int32), ("NHQ", c_int32), ("MDT", c_int32), ("index_real", c_int32),
        ("DT", c_double), ("HM1", c_double), ("HM2", c_double), ("JL", c_double),
        ("H", c_void_p), ("U", c_void_p), ("V", c_void_p), ("Z", c_void_p), ("W", c_void_p),
        ("ZBC", c_void_p), ("ZB1", c_void_p), ("AREA", c_void_p), ("FNC", c_void_p),
        ("BoundaryFeature", c_void_p), ("NHQ1", c_void_p),
        ("ZW", c_void_p), ("QW", c_void_p),
        ("QT", c_void_p), ("DQT", c_void_p), ("ZT", c_void_p), ("DZT", c_void_p),
        ("NAC", c_void_p), ("KLAS", c_void_p),
        ("SIDE", c_void_p), ("COSF", c_void_p), ("SINF", c_void_p), ("SLCOS", c_void_p), ("SLSIN", c_void_p),
        ("FLUX0", c_void_p), ("FLUX1", c_void_p), ("FLUX2", c_void_p), ("FLUX3", c_void_p),
    ]


class Tune(ctypes.Structure):
    """hydro_tune (include/hydro.h)."""

    _fields_ = [
        ("block", c_int32), ("reorder", c_int32), ("xcd_remap", c_int32), ("graph_chunk", c_int32),
        ("persistent_wgs_per_cu", c_int32), ("materialize_flux", c_int32), ("waves_per_simd", c_int32),
        ("scheme", c_int32), ("tile_cells", c_int32), ("spin_limit", c_int32),
        ("persistent_loop", c_int32), ("resident_launch", c_int32), ("regroup_steps", c_int32),
    ]


class BoundaryDesc(ctypes.Structure):
    """hydro_boundary_desc (include/hydro.h)."""

    _fields_ = [
        ("dtype", c_int32), ("CELL", c_int32), ("NDAYS", c_int32), ("rows", c_int32), ("MDT", c_int32),
        ("NHQ", c_int32), ("DT", c_double), ("HM2", c_double),
        ("NZ", c_int32), ("NNZ0", c_int32), ("MBZ", c_void_p), ("NNZ", c_void_p),
        ("NQ", c_int32), ("NNQ0", c_int32), ("MBQ", c_void_p), ("NNQ", c_void_p),
        ("NZQ", c_int32), ("NNZQ0", c_int32), ("MBZQ", c_void_p), ("NNZQ", c_void_p),
        ("ZBC", c_void_p), ("SIDE", c_void_p), ("KLAS", c_void_p),
    ]


def autotune_candidates(hilbert=False):
    """The library's built-in candidate list of hydro_autotune (HILBERT when
    the mesh has cell coordinates)."""
    c = [default_tune(block=256, tile_cells=0), default_tune(block=256, tile_cells=128),
         default_tune(block=128, tile_cells=0), default_tune(block=256, tile_cells=0, waves_per_simd=4),
         default_tune(block=128, scheme=SCHEME_CELL), default_tune(block=256, scheme=SCHEME_CELL),
         default_tune(block=256, reorder=REORDER_RCM)]
    if hilbert:
        c.append(default_tune(block=256, reorder=REORDER_HILBERT))
    return c


def default_tune(**kw) -> Tune:
    t = Tune(block=256, reorder=REORDER_TILES, xcd_remap=1, graph_chunk=16, persistent_wgs_per_cu=0,
             materialize_flux=0, waves_per_simd=0, scheme=SCHEME_TILES, tile_cells=0, spin_limit=0, persistent_loop=0, resident_launch=0,
             regroup_steps=0)
    for k, v in kw.items():
        setattr(t, k, v)
    return t


# Every symbol include/hydro.h declares, with (restype, argtypes).
SIGNATURES = {
    "hydro_mesh_create": (c_int, [POINTER(MeshDesc), POINTER(c_void_p)]),
    "hydro_mesh_destroy": (c_int, [c_void_p]),
    "hydro_prepare": (c_int, [c_void_p, POINTER(Tune)]),
    "hydro_calculate_flux": (c_int, [c_void_p, c_int, c_int, c_void_p]),
    "hydro_update_cell": (c_int, [c_void_p, c_void_p]),
    "hydro_step_n": (c_int, [c_void_p, c_int, c_int, c_void_p]),
    "hydro_set_time": (c_int, [c_void_p, c_int, c_int]),
    "hydro_get_time": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
    "hydro_sync": (c_int, [c_void_p, c_void_p]),
    "hydro_reload": (c_int, [c_void_p, c_void_p]),
    "hydro_last_error": (c_char_p, []),
    "hydro_abi_version": (c_int, []),
    "hydro_dump": (c_int, [c_void_p, c_char_p, c_void_p]),
    "hydro_set_profiling": (c_int, [c_void_p, c_int]),
    "hydro_last_timing": (c_int, [c_void_p, POINTER(c_double), POINTER(c_int), POINTER(c_int)]),
    "hydro_precalculate": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_double] + [c_void_p] * 9),
    "hydro_rcm_order": (c_int, [c_int, c_void_p, c_void_p]),
    "hydro_tile_order": (c_int, [c_int, c_void_p, c_int, c_void_p]),
    # adjoint / tangent-linear support
    "hydro_adj_sizes": (c_int, [c_void_p, POINTER(ctypes.c_size_t), POINTER(ctypes.c_size_t),
                                POINTER(ctypes.c_size_t)]),
    "hydro_state_save": (c_int, [c_void_p, c_void_p, c_void_p]),
    "hydro_state_load": (c_int, [c_void_p, c_void_p, c_void_p]),
    "hydro_step_record": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "hydro_adjoint_step": (c_int, [c_void_p, c_int, c_int] + [c_void_p] * 7),
    "hydro_adjoint_check": (c_int, [c_void_p, POINTER(c_int)]),
    "hydro_tlm_step": (c_int, [c_void_p, c_int, c_int] + [c_void_p] * 7),
    "hydro_gauge_values": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "hydro_gauge_seed": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "hydro_cells_map": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "hydro_states_map": (c_int, [c_void_p] * 6 + [c_int, c_void_p]),
    "hydro_refresh_statics": (c_int, [c_void_p, c_void_p]),
    "hydro_autotune": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, POINTER(Tune), c_void_p, c_void_p]),
    "hydro_set_cell_coords": (c_int, [c_void_p, c_void_p, c_void_p]),
    "hydro_plan_stats": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int), POINTER(c_int), POINTER(c_double)]),
    "hydro_regroup": (c_int, [c_void_p, c_void_p]),
    "hydro_plan_grouping": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int), c_void_p]),
    "hydro_hilbert_order": (c_int, [c_int, c_void_p, c_void_p, c_void_p]),
    "hydro_take_boundary": (c_int, [POINTER(BoundaryDesc), c_char_p] + [c_void_p] * 7),
    "hydro_boundary_increments": (c_int, [c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p]),
    "hydro_misfit_forward": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "hydro_misfit_backward": (c_int, [c_void_p, c_int, c_int] + [c_void_p] * 2 + [c_int] + [c_void_p] * 6
                              + [POINTER(c_int), c_void_p]),
    "hydro_scen_misfit_forward": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_void_p, c_void_p, c_int,
                                          c_void_p, c_void_p]),
    "hydro_scen_misfit_backward": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_void_p, c_void_p, c_int]
                                   + [c_void_p] * 6 + [POINTER(c_int), c_void_p]),
    "hydro_diagnostics": (c_int, [c_void_p, c_void_p, c_void_p]),
    "hydro_diagnostics_noalloc": (c_int, [c_void_p, c_void_p, c_void_p]),
    # batches
    "hydro_batch_create": (c_int, [c_void_p, c_int, POINTER(Tune), POINTER(c_void_p)]),
    "hydro_batch_destroy": (c_int, [c_void_p]),
    "hydro_batch_step_n": (c_int, [c_void_p, c_int, c_int, c_void_p]),
    "hydro_batch_sync": (c_int, [c_void_p, c_void_p]),
}

_lib = None


def _hip_runtimes_mapped():
    """Paths of every libamdhip64 mapped into this process."""
    paths = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64" in line:
                    paths.add(os.path.realpath(line.split()[-1]))
    except OSError:
        pass
    return paths


def load() -> ctypes.CDLL:
    """Load libhydro_amd.so once; raise if it was not built.

    PyTorch-ROCm ships its own libamdhip64 and requests it by a different
    name than the library's DT_NEEDED, so whichever loads first decides
    whether the process ends up with one HIP runtime or two.  Two runtimes
    do not share allocations (our pointer queries would reject torch's
    device memory), so torch is imported first and its runtime serves both.
    """
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C hydro-bench_amd` or __graft_entry__.build()")
    import torch  # noqa: F401  (loads torch's HIP runtime first)

    lib = ctypes.CDLL(LIB_PATH)
    runtimes = _hip_runtimes_mapped()
    if len(runtimes) > 1:
        raise ImportError(f"two HIP runtimes are loaded ({sorted(runtimes)}); import torch before "
                          f"loading libhydro_amd.so")
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.hydro_abi_version() != ABI_VERSION:
        raise ImportError(f"libhydro_amd ABI {lib.hydro_abi_version()} != {ABI_VERSION}")
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != OK:
        msg = load().hydro_last_error()
        raise HydroError(status, msg.decode() if msg else "")
