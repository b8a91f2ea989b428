"""GPU checks of the hand-written adjoint (fp64, variant A).

* one step: the adjoint is the transpose of the tangent-linear model
  (dual numbers through the same step): <lam, J d> = <J^T lam, d> to
  rounding, for random state, Manning and bed directions;
* the whole calibration gradient (24,020-cell floodplain, 500 steps,
  checkpoints every 25, 32 gauges): against fp64 central finite differences
  (h = 1e-6) and the tangent-linear run on random smooth directions, within
  1e-6 relative.
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


def _problem(steps=500, segment=25, n_gauges=32, seed=4004):
    import torch

    from hydro_amd.adjoint import MisfitProblem
    from hydro_amd.device import DeviceMesh

    host = cached_mesh("floodplain_24020")
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


def test_calibration_gradient_vs_fd_and_tlm():
    import torch

    from hydro_amd.adjoint import GaugeMisfit, misfit, tlm_run

    host, prob, n, zbc, rng = _problem(steps=500, segment=25)
    _truth(prob, host, n, zbc, rng)
    nv = n.clone().requires_grad_(True)
    zv = zbc.clone().requires_grad_(True)
    loss = GaugeMisfit.apply(nv, zv, prob)
    loss.backward()
    gn, gz = nv.grad.detach(), zv.grad.detach()
    assert float(loss) > 0
    h = 1e-6
    worst = 0.0
    for i in range(20):
        dn = torch.tensor(_smooth_field(host.xc, host.yc, rng, 600.0, 1.0), device=n.device)
        dz = torch.tensor(_smooth_field(host.xc, host.yc, rng, 600.0, 1.0), device=n.device) if i % 2 else \
            torch.zeros_like(zbc)
        adj = float((gn * dn).sum() + (gz * dz).sum())
        fd = (misfit(prob, n + h * dn, zbc + h * dz) - misfit(prob, n - h * dn, zbc - h * dz)) / (2 * h)
        tl, _ = tlm_run(prob, n, zbc, dn, dz)
        rel_fd = abs(adj - fd) / max(abs(fd), 1e-300)
        rel_tl = abs(adj - tl) / max(abs(tl), 1e-300)
        worst = max(worst, rel_fd)
        print(f"dir {i:2d} adjoint {adj:+.12e} fd {fd:+.12e} tlm {tl:+.12e} rel_fd {rel_fd:.2e} rel_tlm {rel_tl:.2e}",
              flush=True)
        assert rel_tl <= 1e-9
        assert rel_fd <= 1e-6
    prob.dm.close()
