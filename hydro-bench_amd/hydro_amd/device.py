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
                 prepare: bool = True):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("DeviceMesh needs a HIP device; none is visible")
        self.lib = _lib.load()
        self.host = host
        self.device = torch.device("cuda", device)
        self.variant = variant
        self.t = {}
        for k, v in host.arrays.items():
            if k in ("NAP", "XP", "YP", "XYMIN"):
                continue
            self.t[k] = torch.from_numpy(np.ascontiguousarray(v)).to(self.device)
        d = _lib.MeshDesc()
        d.abi_version = _lib.ABI_VERSION
        d.dtype = _lib.F64 if host.dtype == np.float64 else _lib.F32
        d.variant = variant
        d.CELL, d.NDAYS, d.NHQ, d.MDT = host.CELL, host.NDAYS, host.NHQ, host.MDT
        d.DT, d.HM1, d.HM2, d.JL = host.DT, host.HM1, host.HM2, host.JL
        for name, _t in d._fields_:
            if name in self.t:
                setattr(d, name, self.t[name].data_ptr())
        self.desc = d
        h = ctypes.c_void_p()
        torch.cuda.synchronize(self.device)
        _lib.check(self.lib.hydro_mesh_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
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
        cands = list(candidates) if candidates else _lib.autotune_candidates()
        arr = (_lib.Tune * len(cands))(*cands)
        times = (ctypes.c_float * len(cands))()
        best = _lib.Tune()
        _lib.check(self.lib.hydro_autotune(self.h, _lib.MODES[mode], int(steps), arr, len(cands), ctypes.byref(best),
                                           times, self.stream()))
        self.tune = best
        return best, list(zip(cands, list(times)))

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


def run(mesh: HostMesh, steps: Optional[int] = None, mode: str = "two_kernel", variant: int = _lib.VARIANT_A,
        device: int = 0, tune=None):
    """The reference's Python contract (taichi_impl/F2_hydro_taichi_fp64.py:29,773):

    returns ``(step_fn, sync_fn, H, U, V, Z, FLUX0, FLUX1, FLUX2, FLUX3)``.
    ``step_fn(on_step=None)`` runs ``steps`` steps over the (day, kt) grid
    (T2:734-753) and, when ``on_step`` is given, calls ``on_step(s)`` after
    each step once the state has been synchronised (s counts from 1).
    The returned tensors are the live device arrays; in reordered modes they
    are refreshed by ``sync_fn()`` (and before every ``on_step``).
    """
    dm = DeviceMesh(mesh, variant=variant, device=device, tune=tune,
                    prepare=(mode not in ("two_kernel", "graph_two_kernel")))
    total = mesh.max_steps if steps is None else int(steps)
    if total > mesh.max_steps:
        raise _lib.HydroError(_lib.ERR_RANGE, f"{total} steps exceed the NDAYS cap {mesh.max_steps}")

    def sync_fn():
        dm.sync()
        _torch().cuda.synchronize(dm.device)

    def step_fn(on_step: Optional[Callable[[int], None]] = None):
        if on_step is None:
            dm.step(total, mode)
            return
        for s in range(1, total + 1):
            dm.step(1, mode)
            sync_fn()
            on_step(s)

    step_fn.mesh = dm
    t = dm.t
    return step_fn, sync_fn, t["H"], t["U"], t["V"], t["Z"], t["FLUX0"], t["FLUX1"], t["FLUX2"], t["FLUX3"]
