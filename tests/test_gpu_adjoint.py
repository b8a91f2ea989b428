"""GPU checks of the hand-written adjoint (fp64, variant A).

* one step: the adjoint is the transpose of the tangent-linear model
  (dual numbers through the same step): <lam, J d> = <J^T lam, d> to
  rounding, for random state, Manning and bed directions;
* the whole calibration gradient (24,020-cell floodplain, 500 steps,
  checkpoints every 25, 32 gauges) with initial and downstream levels 1.5 m
  higher (every cell under water, so no cell meets the wet/dry thresholds
  HM1/HM2 and the loss is smooth): against fp64 central finite differences (h = 1e-6) and
  the tangent-linear run on random smooth directions, within 1e-6 relative;
* the same gradient from the mesh's own start, with wet/dry fronts: the loss
  is only piecewise smooth there (a cell whose depth crosses HM1 has its
  velocity reset, a jump), so finite differences are not a valid check; the
  adjoint must match the tangent-linear run within 1e-9;
* one step at a developed state with fronts, perturbing state, Manning and
  bed only on cells whose depth and neighbours' depths are far from the
  thresholds: tangent-linear against central differences within 1e-6.
"""
import math

import numpy as np
import pytest

from conftest import cached_mesh

pytestmark = pytest.mark.gpu


def _smooth_field(x, y, rng, length, std, n_modes=24):
    """Seeded smooth random field: a sum of random Fourier modes, scaled to `std`."""
    f = np.zeros_like(x)
    for _ in range(n_modes):
        k = rng.normal(0, 1.0 / length, 2)
        f += math.sqrt(2.0 / n_modes) * np.cos(k[0] * x + k[1] * y + rng.uniform(0, 2 * math.pi))
    return std * f / max(f.std(), 1e-30)


def _problem(steps=500, segment=25, n_gauges=32, seed=4004, deep=False):
    import torch

    from hydro_amd.adjoint import MisfitProblem
    from hydro_amd.device import DeviceMesh

    host = cached_mesh("floodplain_24020", **({"depth_offset": 1.5} if deep else {}))
    rng = np.random.default_rng(seed)
    dm = DeviceMesh(host, device=0)
    dev = dm.device
    n_base = np.sqrt(host.FNC / 9.81)
    level = host.Z.copy()
    level[host.H <= host.HM1] = host.ZBC[host.H <= host.HM1] - 0.0   # dry cells keep their (dry) level
    wet = host.H > 0.5
    gauges = rng.choice(np.nonzero(wet)[0], size=n_gauges, replace=False).astype(np.int32)
    t = lambda a, dt=torch.float64: torch.tensor(a, dtype=dt, device=dev)
    prob = MisfitProblem(dm=dm, level0=t(level), U0=t(host.U), V0=t(host.V), gauges=t(gauges, torch.int32),
                         steps=steps, segment=segment)
    return host, prob, t(n_base), t(host.ZBC.copy()), rng


def _truth(prob, host, n_base, zbc, rng):
    from hydro_amd.adjoint import run_forward

    import torch

    pert = _smooth_field(host.xc, host.yc, rng, 800.0, 0.005)
    n_true = n_base + torch.tensor(pert, device=n_base.device)
    prob.z_obs = torch.zeros((prob.steps, int(prob.gauges.numel())), dtype=torch.float64, device=n_base.device)
    zg, _ = run_forward(prob, n_true, zbc, keep_checkpoints=False)
    prob.z_obs = zg.clone()


def test_adjoint_step_is_transpose_of_tangent_linear():
    import ctypes

    import torch

    import hydro_amd._lib as L

    host, prob, n, zbc, rng = _problem(steps=1)
    dm = prob.dm
    lib, st = dm.lib, dm.stream()
    C = host.CELL
    dev = dm.device
    dm.step(40, "fused")                      # a developed, moving state with fronts
    s0 = torch.empty((C, 4), dtype=torch.float64, device=dev)
    L.check(lib.hydro_state_save(dm.h, ctypes.c_void_p(s0.data_ptr()), st))
    d, kt, spd = dm.get_time()
    jt = d
    states = torch.empty((1, C, 4), dtype=torch.float64, device=dev)
    wsums = torch.empty((1, 3, C), dtype=torch.float64, device=dev)
    L.check(lib.hydro_step_record(dm.h, 1, ctypes.c_void_p(states.data_ptr()), ctypes.c_void_p(wsums.data_ptr()), st))
    g = torch.Generator(device="cpu").manual_seed(5)
    ds = torch.randn((C, 4), generator=g, dtype=torch.float64).to(dev) * 1e-3
    dfnc = torch.randn(C, generator=g, dtype=torch.float64).to(dev) * 1e-4
    dzbc = torch.randn(C, generator=g, dtype=torch.float64).to(dev) * 1e-3
    lam = torch.randn((C, 4), generator=g, dtype=torch.float64).to(dev)
    s1 = torch.empty_like(s0)
    ds1 = torch.empty_like(s0)
    vp = lambda x: ctypes.c_void_p(x.data_ptr())
    L.check(lib.hydro_tlm_step(dm.h, kt, jt, vp(s0), vp(ds), vp(dfnc), vp(dzbc), vp(s1), vp(ds1), st))
    lam_in = torch.empty_like(s0)
    gf = torch.zeros(C, dtype=torch.float64, device=dev)
    gz = torch.zeros(C, dtype=torch.float64, device=dev)
    L.check(lib.hydro_adjoint_step(dm.h, kt, jt, vp(states[0]), vp(wsums[0]), vp(lam), vp(lam_in), vp(gf), vp(gz), st))
    torch.cuda.synchronize()
    # the tangent run's values are the forward step's, bit for bit
    dm_next = torch.empty_like(s0)
    L.check(lib.hydro_state_save(dm.h, vp(dm_next), st))
    torch.cuda.synchronize()
    assert torch.equal(s1, dm_next)
    lhs = float((lam * ds1).sum())
    rhs = float((lam_in * ds).sum() + (gf * dfnc).sum() + (gz * dzbc).sum())
    scale = float((lam.abs() * ds1.abs()).sum()) + 1e-300
    print(f"dot test: lhs {lhs:.15e} rhs {rhs:.15e} rel {abs(lhs - rhs) / scale:.3e}")
    assert abs(lhs - rhs) <= 1e-12 * scale
    dm.close()


def test_misfit_loops_match_per_step_api():
    """hydro_misfit_forward / hydro_misfit_backward (gauge reads inside the
    step kernel, the one-launch adjoint step with gauge seeds) against the
    per-step calls (hydro_gauge_values, hydro_step_record, hydro_gauge_seed,
    the two-kernel hydro_adjoint_step): same bits, with a repeated gauge and
    a last segment shorter than the others."""
    import ctypes

    import torch

    import hydro_amd._lib as L
    from hydro_amd.adjoint import _Buffers, _time_of, run_backward, run_forward

    steps, seg = 60, 8
    host, prob, n, zbc, rng = _problem(steps=steps, segment=seg)
    dm = prob.dm
    lib, st = dm.lib, dm.stream()
    C = host.CELL
    dev = dm.device
    vp = lambda x: ctypes.c_void_p(x.data_ptr())
    g = prob.gauges.clone()
    g[5] = g[2]                                   # a repeated gauge id
    prob.gauges = g
    G = int(g.numel())
    prob.z_obs = torch.tensor(rng.normal(11.0, 0.5, (steps, G)), dtype=torch.float64, device=dev)
    zg, buf = run_forward(prob, n, zbc)
    g_n, g_z = run_backward(prob, buf, zg, n, zbc)
    torch.cuda.synchronize()
    fused_fnc, fused_zbc = buf.g_fnc.clone(), buf.g_zbc.clone()

    # per-step reference
    from hydro_amd.adjoint import _initial_state, _set_params

    _set_params(dm, n, zbc)
    H0, Z0, _ = _initial_state(prob, zbc)
    dm.load_state({"H": H0, "U": prob.U0, "V": prob.V0, "Z": Z0})
    dm.set_time(0, 1)
    nseg = (steps + seg - 1) // seg
    ref = _Buffers(dm, nseg, seg)
    zg_ref = torch.empty_like(zg)
    tmp = torch.empty((C, 4), dtype=torch.float64, device=dev)
    for k in range(steps):
        if k % seg == 0:
            L.check(lib.hydro_state_save(dm.h, vp(ref.ckpt[k // seg]), st))
        dm.step(1, "fused")
        L.check(lib.hydro_state_save(dm.h, vp(tmp), st))
        L.check(lib.hydro_gauge_values(dm.h, vp(tmp), vp(g), G, vp(zg_ref[k]), st))
    torch.cuda.synchronize()
    assert torch.equal(zg, zg_ref)
    assert torch.equal(buf.ckpt[:nseg], ref.ckpt[:nseg])
    resid = (2.0 * (zg_ref - prob.z_obs)).contiguous()
    ref.lam[0].zero_()
    ref.g_fnc.zero_()
    ref.g_zbc.zero_()
    spd = host.steps_per_day
    cur = 0
    for k in reversed(range(nseg)):
        s0 = k * seg
        m = min(seg, steps - s0)
        L.check(lib.hydro_state_load(dm.h, vp(ref.ckpt[k]), st))
        kt, jt = _time_of(s0, spd)
        dm.set_time(jt, kt)
        L.check(lib.hydro_step_record(dm.h, m, vp(ref.states), vp(ref.wsums), st))
        for i in reversed(range(m)):
            L.check(lib.hydro_gauge_seed(dm.h, vp(ref.lam[cur]), vp(g), G, vp(resid[s0 + i]), st))
            kt, jt = _time_of(s0 + i, spd)
            L.check(lib.hydro_adjoint_step(dm.h, kt, jt, vp(ref.states[i]), vp(ref.wsums[i]), vp(ref.lam[cur]),
                                           vp(ref.lam[cur ^ 1]), vp(ref.g_fnc), vp(ref.g_zbc), st))
            cur ^= 1
    torch.cuda.synchronize()
    assert torch.equal(fused_fnc, ref.g_fnc)
    assert torch.equal(fused_zbc, ref.g_zbc)
    assert float(fused_fnc.abs().max()) > 0 and float(fused_zbc.abs().max()) > 0
    dm.close()


def _gradient_check(deep, n_dirs, fd_tol):
    import torch

    from hydro_amd.adjoint import GaugeMisfit, misfit, run_forward, tlm_run

    host, prob, n, zbc, rng = _problem(steps=500, segment=25, deep=deep)
    _truth(prob, host, n, zbc, rng)
    nv = n.clone().requires_grad_(True)
    zv = zbc.clone().requires_grad_(True)
    loss = GaugeMisfit.apply(nv, zv, prob)
    loss.backward()
    gn, gz = nv.grad.detach(), zv.grad.detach()
    assert float(loss.detach()) > 0
    if deep:
        # the checkpoints (every 25 steps) stay far from the thresholds
        _, buf = run_forward(prob, n, zbc)
        nseg = (prob.steps + prob.seg - 1) // prob.seg
        hmin = float(buf.ckpt[:nseg, :, 0].min())
        print(f"deep start: min depth over checkpoints {hmin:.3f} m")
        assert hmin > 10 * host.HM2
    h = 1e-6
    for i in range(n_dirs):
        dn = torch.tensor(_smooth_field(host.xc, host.yc, rng, 600.0, 1.0), device=n.device)
        dz = torch.tensor(_smooth_field(host.xc, host.yc, rng, 600.0, 1.0), device=n.device) if i % 2 else \
            torch.zeros_like(zbc)
        adj = float((gn * dn).sum() + (gz * dz).sum())
        fd = (misfit(prob, n + h * dn, zbc + h * dz) - misfit(prob, n - h * dn, zbc - h * dz)) / (2 * h)
        tl, _ = tlm_run(prob, n, zbc, dn, dz)
        rel_fd = abs(adj - fd) / max(abs(fd), 1e-300)
        rel_tl = abs(adj - tl) / max(abs(tl), 1e-300)
        print(f"dir {i:2d} adjoint {adj:+.12e} fd {fd:+.12e} tlm {tl:+.12e} rel_fd {rel_fd:.2e} rel_tlm {rel_tl:.2e}",
              flush=True)
        assert rel_tl <= 1e-9
        if fd_tol is not None:
            assert rel_fd <= fd_tol
    prob.dm.close()


def test_calibration_gradient_vs_fd_and_tlm_deep_start():
    _gradient_check(deep=True, n_dirs=20, fd_tol=1e-6)   # C4: 20 random directions, h = 1e-6


def test_calibration_gradient_vs_tlm_with_fronts():
    _gradient_check(deep=False, n_dirs=4, fd_tol=None)


def test_one_step_fd_at_fronts():
    import ctypes

    import torch

    import hydro_amd._lib as L

    host, prob, n, zbc, rng = _problem(steps=1)
    dm = prob.dm
    lib, st = dm.lib, dm.stream()
    C = host.CELL
    dev = dm.device
    vp = lambda x: ctypes.c_void_p(x.data_ptr())
    dm.step(40, "fused")
    d, kt, _ = dm.get_time()
    jt = d
    s0 = torch.empty((C, 4), dtype=torch.float64, device=dev)
    L.check(lib.hydro_state_save(dm.h, vp(s0), st))
    ext = [torch.empty(C, dtype=torch.float64, device=dev) for _ in range(4)]
    L.check(lib.hydro_states_map(dm.h, vp(s0), *[vp(x) for x in ext], 1, st))
    torch.cuda.synchronize()
    H = ext[0].cpu().numpy()
    # cells far from the thresholds, with all neighbours far from them too
    deep = H > 0.1
    nac = host.NAC.reshape(C, 4)
    nb_ok = np.ones(C, dtype=bool)
    for j in range(4):
        nbr = nac[:, j]
        has = nbr > 0
        nb_ok[has] &= deep[nbr[has] - 1]
    mask = deep & nb_ok
    print(f"perturbed cells: {int(mask.sum())} of {C}")
    assert mask.sum() > C // 10
    g = np.random.default_rng(11)
    m_t = torch.tensor(mask.astype(np.float64), device=dev)
    rnd = lambda scale: torch.tensor(g.normal(0, scale, C), device=dev) * m_t
    dH, dU, dV, dZ = rnd(1e-2), rnd(1e-2), rnd(1e-2), rnd(1e-2)
    ds = torch.empty((C, 4), dtype=torch.float64, device=dev)
    L.check(lib.hydro_states_map(dm.h, vp(ds), vp(dH), vp(dU), vp(dV), vp(dZ), 0, st))
    dfnc_ext = rnd(1e-3)
    dzbc_ext = rnd(1e-2)
    dfnc = torch.empty(C, dtype=torch.float64, device=dev)
    dzbc = torch.empty(C, dtype=torch.float64, device=dev)
    L.check(lib.hydro_cells_map(dm.h, vp(dfnc_ext), vp(dfnc), 0, st))
    L.check(lib.hydro_cells_map(dm.h, vp(dzbc_ext), vp(dzbc), 0, st))
    lam = torch.tensor(g.normal(0, 1, (C, 4)), device=dev)
    fnc0 = dm.t["FNC"].clone()
    zbc0 = dm.t["ZBC"].clone()

    def stepped(sgn, h):
        with torch.no_grad():
            dm.t["FNC"].copy_(fnc0 + sgn * h * dfnc_ext)
            dm.t["ZBC"].copy_(zbc0 + sgn * h * dzbc_ext)
            dm.t["ZB1"].copy_(dm.t["ZBC"] + host.HM1)
        L.check(lib.hydro_refresh_statics(dm.h, st))
        sp = s0 + sgn * h * ds
        L.check(lib.hydro_state_load(dm.h, vp(sp), st))
        dm.set_time(jt, kt)
        dm.step(1, "fused")
        s1 = torch.empty_like(s0)
        L.check(lib.hydro_state_save(dm.h, vp(s1), st))
        torch.cuda.synchronize()
        return s1

    # tangent-linear at the base point (statics restored)
    stepped(0.0, 0.0)
    s1 = torch.empty_like(s0)
    ds1 = torch.empty_like(s0)
    L.check(lib.hydro_tlm_step(dm.h, kt, jt, vp(s0), vp(ds), vp(dfnc), vp(dzbc), vp(s1), vp(ds1), st))
    torch.cuda.synchronize()
    tl = float((lam * ds1).sum())
    for h in (1e-5, 1e-6):
        fd_vec = (stepped(1.0, h) - stepped(-1.0, h)) / (2 * h)
        fd = float((lam * fd_vec).sum())
        rel = abs(fd - tl) / max(abs(tl), 1e-300)
        print(f"one step h={h:.0e}: tlm {tl:+.12e} fd {fd:+.12e} rel {rel:.2e}", flush=True)
        assert rel <= 1e-6
    dm.close()
