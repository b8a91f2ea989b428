"""Device meshes and the reference's Python ``run()``/``step_fn`` contract.

PyTorch-ROCm provides device memory and the HIP stream; every kernel is the
library's (libhydro_amd.so through the C ABI).  There is no CPU execution
path: constructing a :class:`DeviceMesh` without a GPU or without the built
library raises.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Optional

import numpy as np

from . import _lib
from .mesh import HostMesh

_TORCH_DTYPES = None


def _torch():
    import torch

    return torch


def _tdtype(np_dtype):
    torch = _torch()
    return {np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32,
            np.dtype(np.int32): torch.int32}[np.dtype(np_dtype)]


class DeviceMesh:
    """A HostMesh uploaded to ``cuda:<device>`` plus its hydro_mesh handle.

    ``self.t[name]`` are torch tensors in the reference layout; H, U, V, Z,
    W, KLAS and FLUX0..3 are live (written by the kernels).  After steps in
    a reordered mode (fused / graph / persistent) call :meth:`sync` before
    reading them.
    """

    def __init__(self, host: HostMesh, variant: int = _lib.VARIANT_A, device: int = 0, tune=None,
                 prepare: bool = True, index_real: int = 0):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("DeviceMesh needs a HIP device; none is visible")
        self.lib = _lib.load()
        self.host = host
        self.device = torch.device("cuda", device)
        self.variant = variant
        self.t = {}
        real_idx = {"NAC": bool(index_real & _lib.NAC_REAL), "KLAS": bool(index_real & _lib.KLAS_REAL)}
        for k, v in host.arrays.items():
            if k in ("NAP", "XP", "YP", "XYMIN"):
                continue
            if real_idx.get(k):   # NAC / KLAS as Real, as the CUDA SideView holds them
                v = v.astype(host.dtype)
            self.t[k] = torch.from_numpy(np.ascontiguousarray(v)).to(self.device)
        d = _lib.MeshDesc()
        d.abi_version = _lib.ABI_VERSION
        d.dtype = _lib.F64 if host.dtype == np.float64 else _lib.F32
        d.variant = variant
        d.CELL, d.NDAYS, d.NHQ, d.MDT = host.CELL, host.NDAYS, host.NHQ, host.MDT
        d.index_real = int(index_real)
        d.DT, d.HM1, d.HM2, d.JL = host.DT, host.HM1, host.HM2, host.JL
        for name, _t in d._fields_:
            if name in self.t:
                setattr(d, name, self.t[name].data_ptr())
        self.desc = d
        h = ctypes.c_void_p()
        torch.cuda.synchronize(self.device)
        _lib.check(self.lib.hydro_mesh_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        if host.xc is not None and host.yc is not None:   # for the HILBERT cell order
            xc = np.ascontiguousarray(host.xc, dtype=np.float64)
            yc = np.ascontiguousarray(host.yc, dtype=np.float64)
            _lib.check(self.lib.hydro_set_cell_coords(h, xc.ctypes.data_as(ctypes.c_void_p),
                                                      yc.ctypes.data_as(ctypes.c_void_p)))
        if prepare:
            self.prepare(tune)

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.hydro_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare(self, tune=None):
        t = tune if tune is not None else _lib.default_tune()
        _lib.check(self.lib.hydro_prepare(self.h, ctypes.byref(t)))

    # -- stream -------------------------------------------------------------
    def autotune(self, mode: str = "fused", steps: int = 50, candidates=None):
        """Time candidate tunings of the reordered path (hydro_autotune) and
        keep the fastest; the state and time are left as they were.  Returns
        (best Tune, [(Tune, us per step), ...]); -1 marks a candidate that
        could not run."""
        cands = list(candidates) if candidates else _lib.autotune_candidates(hilbert=self.host.xc is not None)
        arr = (_lib.Tune * len(cands))(*cands)
        times = (ctypes.c_float * len(cands))()
        best = _lib.Tune()
        _lib.check(self.lib.hydro_autotune(self.h, _lib.MODES[mode], int(steps), arr, len(cands), ctypes.byref(best),
                                           times, self.stream()))
        self.tune = best
        return best, list(zip(cands, list(times)))

    def plan_stats(self):
        """(tiles, pair tasks, one-slot tasks, edges per thread) of the prepared plan."""
        t, p, s, e = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _lib.check(self.lib.hydro_plan_stats(self.h, ctypes.byref(t), ctypes.byref(p), ctypes.byref(s),
                                             ctypes.byref(e)))
        return t.value, p.value, s.value, e.value

    def regroup(self):
        """Regroup the plan's tasks by the current wet/dry state (hydro_regroup)."""
        _lib.check(self.lib.hydro_regroup(self.h, self.stream()))

    def plan_grouping(self):
        """(pair chunks, pair chunks mixing wet/dry classes in the current state)."""
        pc, mc = ctypes.c_int(), ctypes.c_int()
        _lib.check(self.lib.hydro_plan_grouping(self.h, ctypes.byref(pc), ctypes.byref(mc), self.stream()))
        return pc.value, mc.value

    def stream(self):
        return ctypes.c_void_p(_torch().cuda.current_stream(self.device).cuda_stream)

    # -- the reference operator API -----------------------------------------
    def calculate_flux(self, kt: int, jt: int):
        _lib.check(self.lib.hydro_calculate_flux(self.h, kt, jt, self.stream()))

    def update_cell(self):
        _lib.check(self.lib.hydro_update_cell(self.h, self.stream()))

    def step(self, n: int = 1, mode: str = "two_kernel"):
        _lib.check(self.lib.hydro_step_n(self.h, int(n), _lib.MODES[mode], self.stream()))

    def set_time(self, day: int, kt: int):
        _lib.check(self.lib.hydro_set_time(self.h, day, kt))

    def get_time(self):
        d, k, s = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(self.lib.hydro_get_time(self.h, ctypes.byref(d), ctypes.byref(k), ctypes.byref(s)))
        return d.value, k.value, s.value

    def sync(self):
        _lib.check(self.lib.hydro_sync(self.h, self.stream()))

    def reload(self):
        _lib.check(self.lib.hydro_reload(self.h, self.stream()))

    def dump(self, path: str):
        _lib.check(self.lib.hydro_dump(self.h, path.encode(), self.stream()))

    def diagnostics(self, out=None):
        """Device-reduced conservation and health figures of the current state
        (hydro_diagnostics): a dict keyed by the reference checker's names.
        With `out` (a float64 device tensor of len(DIAG_FIELDS)) the values
        stay on the device and `out` is returned instead."""
        torch = _torch()
        buf = out if out is not None else torch.empty(len(_lib.DIAG_FIELDS), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.hydro_diagnostics(self.h, ctypes.c_void_p(buf.data_ptr()), self.stream()))
        if out is not None:
            return out
        vals = buf.cpu().tolist()
        return dict(zip(_lib.DIAG_FIELDS, vals))

    def set_profiling(self, on: bool):
        _lib.check(self.lib.hydro_set_profiling(self.h, 1 if on else 0))

    def last_timing(self):
        ms, la, st = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        _lib.check(self.lib.hydro_last_timing(self.h, ctypes.byref(ms), ctypes.byref(la), ctypes.byref(st)))
        return ms.value, la.value, st.value

    # -- state helpers ----------------------------------------------------------
    def load_state(self, state: dict):
        """Overwrite H U V Z (and W) from host arrays or tensors, then re-gather."""
        torch = _torch()
        for k, v in state.items():
            if isinstance(v, torch.Tensor):
                self.t[k].copy_(v)
            else:
                self.t[k].copy_(torch.from_numpy(np.ascontiguousarray(v)).to(self.device))
        self.reload()

    def fetch(self, *names):
        self.sync()
        _torch().cuda.synchronize(self.device)
        return {k: self.t[k].cpu().numpy() for k in names}


# Mesh names of the reference's Taichi loader (taichi_impl/mesh_loader_f2.py:61-90:
# "default" is its 24,020-cell mesh, "20w" its 207,234-cell one), mapped to the
# seeded synthetic meshes of the same sizes; a directory is read as a .DAT mesh.
MESH_NAMES = {"default": "floodplain_24020", "20w": "floodplain_207234"}
BACKENDS = ("cuda", "hip", "rocm", "gpu")


def _mesh_of(mesh, dtype):
    import os

    from . import mesh as M

    if isinstance(mesh, HostMesh):
        return mesh
    if mesh in MESH_NAMES:
        return M.make_config(MESH_NAMES[mesh], dtype=dtype)
    if mesh in M.CONFIGS:
        return M.make_config(mesh, dtype=dtype)
    if isinstance(mesh, str) and os.path.isdir(mesh):
        from .dat import load_dat

        return load_dat(mesh, dtype=dtype)
    raise ValueError(f"unknown mesh {mesh!r}: a HostMesh, one of {sorted(MESH_NAMES) + sorted(M.CONFIGS)}, "
                     f"or a .DAT directory")


def run(days: int = 10, backend: str = "cuda", mesh="default", steps: Optional[int] = None, *,
        mode: str = "two_kernel", variant: int = _lib.VARIANT_A, dtype=np.float64, device: int = 0, tune=None,
        index_real: int = 0):
    """The reference's Python entry point, ``run(days, backend, mesh, steps)``
    (taichi_impl/F2_hydro_taichi_fp64.py:29, step loop :734-753, return :773):

    returns ``(step_fn, sync_fn, H, U, V, Z, FLUX0, FLUX1, FLUX2, FLUX3)``.

    * ``backend``: "cuda" (or "hip", "rocm", "gpu") runs on ``cuda:<device>``;
      there is no CPU backend (the reference's "cpu" runs Taichi on the host).
    * ``mesh``: a HostMesh, "default" / "20w" (the reference loader's names,
      as seeded synthetic meshes of its 24,020 and 207,234 cells), a workload
      name of hydro_amd.CONFIGS, or a directory in the reference's .DAT format.
    * ``step_fn(on_step=None)`` runs, from day 0, the steps of
      ``min(days, NDAYS)`` days (``steps_per_day - 1`` each) or, when
      ``steps`` is given, ``steps`` steps over up to NDAYS days, stopping at the
      NDAYS cap as the reference's loop does; with ``on_step``, the state is
      synchronised after every step and ``on_step(s)`` called (s from 1).
      Every call starts again at day 0, kt 1, from the state it finds.
    * ``sync_fn()`` makes the returned tensors current (reordered modes keep
      the state in the library's layout until then) and waits for the device.

    The returned tensors are the live device arrays in the reference layout.
    ``mode`` picks the execution mode (the reference runs the two-kernel step).
    """
    if backend not in BACKENDS:
        raise ValueError(f"backend {backend!r}: this library runs on the GPU only ({', '.join(BACKENDS)})")
    host = _mesh_of(mesh, dtype)
    dm = DeviceMesh(host, variant=variant, device=device, tune=tune,
                    prepare=(mode not in ("two_kernel", "graph_two_kernel")), index_real=index_real)
    per_day = max(host.steps_per_day - 1, 1)
    total = int(steps) if steps is not None else per_day * int(days)
    day_limit = host.NDAYS if steps is not None else min(int(days), host.NDAYS)
    total = max(0, min(total, day_limit * per_day))

    def sync_fn():
        dm.sync()
        _torch().cuda.synchronize(dm.device)

    def step_fn(on_step: Optional[Callable[[int], None]] = None):
        dm.set_time(0, 1)
        if on_step is None:
            dm.step(total, mode)
            return
        for s in range(1, total + 1):
            dm.step(1, mode)
            sync_fn()
            on_step(s)

    step_fn.mesh = dm
    step_fn.total_steps = total
    t = dm.t
    return step_fn, sync_fn, t["H"], t["U"], t["V"], t["Z"], t["FLUX0"], t["FLUX1"], t["FLUX2"], t["FLUX3"]
