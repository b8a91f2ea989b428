"""Gradients of a gauge-level misfit with respect to per-cell Manning n and
bed elevation, through the library's hand-written adjoint kernels.

``GaugeMisfit`` is a ``torch.autograd.Function``:

    loss = GaugeMisfit.apply(n, zbc, problem)

with ``n`` and ``zbc`` float64 tensors on the device in the caller's cell
order.  The mesh may be fp64 or fp32: an fp32 mesh runs the forward steps,
the checkpoints and the recomputation in fp32, and the adjoint steps in
double on the recorded fp32 states (the gradient of the step's arithmetic
along the fp32 trajectory), accumulating the gradients in double.  The loss
is

    L = sum_{steps s} sum_{gauges g} (Z_g^s - Zobs_g^s)^2

over the state after every step.  The forward pass runs the fused step and
keeps a checkpoint of the state every ``segment`` steps (sqrt(T) by default);
the backward pass replays each segment from its checkpoint (recording the
step inputs and flux sums) and then runs the adjoint steps backwards through
it, so memory is O(sqrt(T)) states instead of O(T).

The initial condition is a fixed water level: depth and level follow
preCalculate's rule (mesh.cpp:350-361) from the level and the bed, so the bed
gradient includes the initial-depth dependence.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from .device import DeviceMesh
from .mesh import HostMesh


def _torch():
    import torch

    return torch


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


@dataclass
class MisfitProblem:
    """A mesh, a start state, gauges and the observations to fit."""

    dm: DeviceMesh
    level0: "object"            # torch f64 [C]: initial water level (caller order)
    U0: "object"
    V0: "object"
    gauges: "object"            # torch int32 [G] original cell ids
    steps: int
    z_obs: Optional["object"] = None   # torch f64 [steps, G]
    segment: int = 0            # checkpoint interval; 0 = ceil(sqrt(steps))
    stats: dict = field(default_factory=dict)

    @property
    def seg(self):
        return self.segment if self.segment > 0 else max(1, int(math.ceil(math.sqrt(self.steps))))


def _time_of(g, spd):
    per = spd - 1
    return 1 + g % per, g // per   # (kt, jt)


class _Buffers:
    def __init__(self, dm: DeviceMesh, nseg: int, seg: int):
        torch = _torch()
        lib = dm.lib
        sb, wb, cb = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        _lib.check(lib.hydro_adj_sizes(dm.h, ctypes.byref(sb), ctypes.byref(wb), ctypes.byref(cb)))
        C = dm.host.CELL
        dev = dm.device
        self.C = C
        # the trajectory (checkpoints, recorded states and flux sums) in the
        # mesh's type; adjoint states and gradients in double
        tt = torch.float64 if dm.host.dtype == np.float64 else torch.float32
        self.ckpt = torch.empty((nseg + 1, C, 4), dtype=tt, device=dev)
        # two segments of records: one is recomputed while the other's
        # adjoint steps run
        self.states = torch.empty((2 * seg, C, 4), dtype=tt, device=dev)
        self.wsums = torch.empty((2 * seg, 3, C), dtype=tt, device=dev)
        self.lam = torch.zeros((2, C, 4), dtype=torch.float64, device=dev)
        self.g_fnc = torch.zeros(C, dtype=torch.float64, device=dev)
        self.g_zbc = torch.zeros(C, dtype=torch.float64, device=dev)
        self.scratch = torch.empty((C, 4), dtype=tt, device=dev)
        rs = self.ckpt.element_size()
        assert sb.value == C * 4 * rs and wb.value == C * 3 * rs


def _set_params(dm: DeviceMesh, n, zbc):
    """Write FNC = 9.81 n^2, ZBC and ZB1 = ZBC + HM1 into the mesh and refresh."""
    torch = _torch()
    with torch.no_grad():
        dm.t["FNC"].copy_(9.81 * n * n)
        zt = zbc.to(dm.t["ZBC"].dtype)   # ZB1 = ZBC + HM1 in the mesh's type, as the steps form it
        dm.t["ZBC"].copy_(zt)
        dm.t["ZB1"].copy_(zt + torch.tensor(dm.host.HM1, dtype=zt.dtype, device=zt.device))
    _lib.check(dm.lib.hydro_refresh_statics(dm.h, dm.stream()))


def _initial_state(prob: MisfitProblem, zbc):
    """preCalculate's dry/wet rule from the fixed level (mesh.cpp:354-361)."""
    torch = _torch()
    HM1 = prob.dm.host.HM1
    lvl = prob.level0
    wet = lvl > zbc
    H = torch.where(wet, lvl - zbc, torch.full_like(lvl, HM1))
    Z = torch.where(wet, lvl, zbc + HM1)
    return H, Z, wet


def run_forward(prob: MisfitProblem, n, zbc, keep_checkpoints=True):
    """Forward run (hydro_misfit_forward: one fused step and one gauge read per
    step, a checkpoint every ``seg`` steps); returns gauge Z [steps, G] and the
    buffers holding the checkpoints."""
    torch = _torch()
    dm = prob.dm
    lib = dm.lib
    st = dm.stream()
    _set_params(dm, n, zbc)
    H0, Z0, _ = _initial_state(prob, zbc)
    dm.load_state({"H": H0, "U": prob.U0, "V": prob.V0, "Z": Z0})
    dm.set_time(0, 1)
    seg = prob.seg
    nseg = (prob.steps + seg - 1) // seg
    buf = _Buffers(dm, nseg, seg) if keep_checkpoints else None
    G = int(prob.gauges.numel())
    zg = torch.empty((prob.steps, G), dtype=torch.float64, device=dm.device)
    ckpt = _vp(buf.ckpt) if buf is not None else ctypes.c_void_p()
    _lib.check(lib.hydro_misfit_forward(dm.h, prob.steps, seg, ckpt, _vp(prob.gauges), G, _vp(zg), st))
    return zg, buf


def run_backward(prob: MisfitProblem, buf: _Buffers, zg, n, zbc, gL=1.0):
    """Adjoint sweep (hydro_misfit_backward: per segment, restore the checkpoint,
    re-run the segment recording step inputs and flux sums, then the adjoint
    steps backwards); returns (dL/dn, dL/dzbc) in the caller's cell order."""
    torch = _torch()
    dm = prob.dm
    lib = dm.lib
    st = dm.stream()
    seg = prob.seg
    G = int(prob.gauges.numel())
    resid = (2.0 * gL * (zg - prob.z_obs)).contiguous()    # dL/dZ at the gauges, per step
    buf.g_fnc.zero_()
    buf.g_zbc.zero_()
    dm.set_time(0, 1)
    idx = ctypes.c_int()
    _lib.check(lib.hydro_misfit_backward(dm.h, prob.steps, seg, _vp(buf.ckpt), _vp(prob.gauges), G, _vp(resid),
                                         _vp(buf.states), _vp(buf.wsums), _vp(buf.lam), _vp(buf.g_fnc),
                                         _vp(buf.g_zbc), ctypes.byref(idx), st))
    cur = idx.value
    bad = ctypes.c_int()
    _lib.check(lib.hydro_adjoint_check(dm.h, ctypes.byref(bad)))
    if bad.value:
        raise _lib.HydroError(_lib.ERR_UNSUPPORTED, "the mesh has slot kinds without an adjoint")
    C = dm.host.CELL
    g_fnc = torch.empty(C, dtype=torch.float64, device=dm.device)
    g_zbc = torch.empty(C, dtype=torch.float64, device=dm.device)
    lamH, lamU, lamV, lamZ = (torch.empty(C, dtype=torch.float64, device=dm.device) for _ in range(4))
    _lib.check(lib.hydro_cells_map(dm.h, _vp(buf.g_fnc), _vp(g_fnc), 1, st))
    _lib.check(lib.hydro_cells_map(dm.h, _vp(buf.g_zbc), _vp(g_zbc), 1, st))
    _lib.check(lib.hydro_states_map(dm.h, _vp(buf.lam[cur]), _vp(lamH), _vp(lamU), _vp(lamV), _vp(lamZ), 1, st))
    # initial condition: wet H0 = level - zbc, dry Z0 = zbc + HM1
    _, _, wet = _initial_state(prob, zbc)
    g_zbc = g_zbc + torch.where(wet, -lamH, lamZ)
    g_n = g_fnc * 2.0 * 9.81 * n
    return g_n, g_zbc


class _ScenBuffers:
    """Buffers of the scenario loops: each record holds all S scenarios."""

    def __init__(self, dm: DeviceMesh, S: int, nseg: int, seg: int):
        torch = _torch()
        C = dm.host.CELL
        dev = dm.device
        tt = torch.float64 if dm.host.dtype == np.float64 else torch.float32
        self.S, self.C = S, C
        self.ckpt = torch.empty((nseg + 1, S, C, 4), dtype=tt, device=dev)
        self.states = torch.empty((2 * seg, S, C, 4), dtype=tt, device=dev)
        self.wsums = torch.empty((2 * seg, S, 3, C), dtype=tt, device=dev)
        self.lam = torch.zeros((2, S, C, 4), dtype=torch.float64, device=dev)
        self.g_fnc = torch.zeros((S, C), dtype=torch.float64, device=dev)
        self.g_zbc = torch.zeros((S, C), dtype=torch.float64, device=dev)


def run_forward_scen(prob: MisfitProblem, ns, zbc, keep_checkpoints=True):
    """The forward run of S Manning fields at once (``ns`` [S, C], caller
    order; hydro_scen_misfit_forward: one launch per step for all of them);
    returns gauge Z [steps, S, G] and the buffers holding the checkpoints."""
    torch = _torch()
    dm = prob.dm
    st = dm.stream()
    _set_params(dm, ns[0], zbc)
    H0, Z0, _ = _initial_state(prob, zbc)
    dm.load_state({"H": H0, "U": prob.U0, "V": prob.V0, "Z": Z0})
    dm.set_time(0, 1)
    S = int(ns.shape[0])
    seg = prob.seg
    nseg = (prob.steps + seg - 1) // seg
    buf = _ScenBuffers(dm, S, nseg, seg) if keep_checkpoints else None
    G = int(prob.gauges.numel())
    fnc = (9.81 * ns * ns).contiguous()
    zg = torch.empty((prob.steps, S, G), dtype=torch.float64, device=dm.device)
    ckpt = _vp(buf.ckpt) if buf is not None else ctypes.c_void_p()
    _lib.check(dm.lib.hydro_scen_misfit_forward(dm.h, S, _vp(fnc), prob.steps, seg, ckpt, _vp(prob.gauges), G,
                                                _vp(zg), st))
    return zg, buf


def run_backward_scen(prob: MisfitProblem, buf: _ScenBuffers, zg, ns, zbc, gL=1.0):
    """Adjoint sweep of run_forward_scen for all S fields at once (one launch
    per adjoint step); returns (dL_s/dn_s, dL_s/dzbc) [S, C], caller order,
    where L_s is scenario s's misfit."""
    torch = _torch()
    dm = prob.dm
    lib = dm.lib
    st = dm.stream()
    S, C = buf.S, buf.C
    G = int(prob.gauges.numel())
    resid = (2.0 * gL * (zg - prob.z_obs[:, None, :])).contiguous()
    fnc = (9.81 * ns * ns).contiguous()
    buf.g_fnc.zero_()
    buf.g_zbc.zero_()
    idx = ctypes.c_int()
    _lib.check(lib.hydro_scen_misfit_backward(dm.h, S, _vp(fnc), prob.steps, prob.seg, _vp(buf.ckpt),
                                              _vp(prob.gauges), G, _vp(resid), _vp(buf.states), _vp(buf.wsums),
                                              _vp(buf.lam), _vp(buf.g_fnc), _vp(buf.g_zbc), ctypes.byref(idx), st))
    cur = idx.value
    bad = ctypes.c_int()
    _lib.check(lib.hydro_adjoint_check(dm.h, ctypes.byref(bad)))
    if bad.value:
        raise _lib.HydroError(_lib.ERR_UNSUPPORTED, "the mesh has slot kinds without an adjoint")
    _, _, wet = _initial_state(prob, zbc)
    g_n = torch.empty((S, C), dtype=torch.float64, device=dm.device)
    g_z = torch.empty_like(g_n)
    lamH, lamU, lamV, lamZ = (torch.empty(C, dtype=torch.float64, device=dm.device) for _ in range(4))
    for k in range(S):
        _lib.check(lib.hydro_cells_map(dm.h, _vp(buf.g_fnc[k]), _vp(g_n[k]), 1, st))
        _lib.check(lib.hydro_cells_map(dm.h, _vp(buf.g_zbc[k]), _vp(g_z[k]), 1, st))
        lam_k = buf.lam[cur, k].contiguous()
        _lib.check(lib.hydro_states_map(dm.h, _vp(lam_k), _vp(lamH), _vp(lamU), _vp(lamV), _vp(lamZ), 1, st))
        g_z[k] += torch.where(wet, -lamH, lamZ)
    g_n = g_n * 2.0 * 9.81 * ns
    return g_n, g_z


class GaugeMisfit:
    """torch.autograd.Function wrapper (built lazily so importing needs no torch)."""

    _fn = None

    @classmethod
    def apply(cls, n, zbc, prob: MisfitProblem):
        if cls._fn is None:
            torch = _torch()

            class _F(torch.autograd.Function):
                @staticmethod
                def forward(ctx, n_, zbc_, prob_):
                    zg, buf = run_forward(prob_, n_.detach(), zbc_.detach())
                    ctx.prob, ctx.buf, ctx.zg = prob_, buf, zg
                    ctx.save_for_backward(n_, zbc_)
                    return ((zg - prob_.z_obs) ** 2).sum()

                @staticmethod
                def backward(ctx, gL):
                    n_, zbc_ = ctx.saved_tensors
                    g_n, g_z = run_backward(ctx.prob, ctx.buf, ctx.zg, n_.detach(), zbc_.detach(), float(gL))
                    return g_n, g_z, None

            cls._fn = _F
        return cls._fn.apply(n, zbc, prob)


def misfit(prob: MisfitProblem, n, zbc):
    """Loss only (no checkpoints kept)."""
    zg, _ = run_forward(prob, n, zbc, keep_checkpoints=False)
    return float(((zg - prob.z_obs) ** 2).sum())


def tlm_run(prob: MisfitProblem, n, zbc, dn, dzbc):
    """Directional derivative of the loss by the tangent-linear model (dual
    numbers through every step), for checking the adjoint."""
    torch = _torch()
    dm = prob.dm
    lib = dm.lib
    st = dm.stream()
    spd = dm.host.steps_per_day
    C = dm.host.CELL
    dev = dm.device
    _set_params(dm, n, zbc)
    H0, Z0, wet = _initial_state(prob, zbc)
    dH0 = torch.where(wet, -dzbc, torch.zeros_like(dzbc))
    dZ0 = torch.where(wet, torch.zeros_like(dzbc), dzbc)
    dm.load_state({"H": H0, "U": prob.U0, "V": prob.V0, "Z": Z0})
    dm.set_time(0, 1)
    s_a = torch.empty((C, 4), dtype=torch.float64, device=dev)
    s_b = torch.empty_like(s_a)
    ds_a = torch.empty_like(s_a)
    ds_b = torch.empty_like(s_a)
    _lib.check(lib.hydro_state_save(dm.h, _vp(s_a), st))
    zeros = torch.zeros(C, dtype=torch.float64, device=dev)
    _lib.check(lib.hydro_states_map(dm.h, _vp(ds_a), _vp(dH0), _vp(zeros), _vp(zeros), _vp(dZ0), 0, st))
    dfnc_ext = 2.0 * 9.81 * n * dn
    dfnc = torch.empty(C, dtype=torch.float64, device=dev)
    dz = torch.empty(C, dtype=torch.float64, device=dev)
    _lib.check(lib.hydro_cells_map(dm.h, _vp(dfnc_ext), _vp(dfnc), 0, st))
    _lib.check(lib.hydro_cells_map(dm.h, _vp(dzbc.contiguous()), _vp(dz), 0, st))
    G = int(prob.gauges.numel())
    zg = torch.empty((prob.steps, G), dtype=torch.float64, device=dev)
    dzg = torch.empty_like(zg)
    for s in range(prob.steps):
        kt, jt = _time_of(s, spd)
        _lib.check(lib.hydro_tlm_step(dm.h, kt, jt, _vp(s_a), _vp(ds_a), _vp(dfnc), _vp(dz), _vp(s_b), _vp(ds_b), st))
        _lib.check(lib.hydro_gauge_values(dm.h, _vp(s_b), _vp(prob.gauges), G, _vp(zg[s]), st))
        _lib.check(lib.hydro_gauge_values(dm.h, _vp(ds_b), _vp(prob.gauges), G, _vp(dzg[s]), st))
        s_a, s_b = s_b, s_a
        ds_a, ds_b = ds_b, ds_a
    dL = float((2.0 * (zg - prob.z_obs) * dzg).sum())
    return dL, zg
