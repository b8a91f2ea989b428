"""GPU: batched Manning calibration at its configured size (the 207,234-cell
floodplain, 16 hypotheses, 900 steps, checkpoints every 30, 20 gradient
steps each, fp32 trajectory with fp64 adjoint and gradients, all hypotheses
in one scenario batch).

Every hypothesis lowers its fp64 misfit; for two of them (the first and the
one with the median reduction) the misfit of the final field reported by the
GPU's fp64 forward run equals the fp64 CPU oracle's for the same field.
"""
import time

import numpy as np
import pytest

from conftest import cached_mesh
from test_gpu_calibrate import _oracle_misfit

pytestmark = pytest.mark.gpu


def test_c5_calibration_at_size(oracle):
    import torch

    from hydro_amd.adjoint import MisfitProblem, run_forward
    from hydro_amd.calibrate import calibrate_many, smooth_field
    from hydro_amd.device import DeviceMesh

    host = cached_mesh("floodplain_207234")
    h32 = cached_mesh("floodplain_207234", dtype=np.float32)
    steps, S, iters = 900, 16, 20
    assert host.CELL == 207234 and host.max_steps >= steps
    dm64 = DeviceMesh(host, device=0)
    dev = dm64.device
    t = lambda x, dt=torch.float64: torch.tensor(x, dtype=dt, device=dev)
    rng = np.random.default_rng(5005)
    level = host.Z.copy()
    level[host.H <= host.HM1] = host.ZBC[host.H <= host.HM1]
    gauges = rng.choice(np.nonzero(host.H > 0.5)[0], size=64, replace=False).astype(np.int32)

    def problem(dm):
        return MisfitProblem(dm=dm, level0=t(level), U0=t(host.U), V0=t(host.V), gauges=t(gauges, torch.int32),
                             steps=steps)

    p64 = problem(dm64)
    zbc = t(host.ZBC.copy())
    field = lambda seed: np.clip(smooth_field(host.xc, host.yc, seed, 0.04, 0.01, 2000.0), 0.005, 0.2)
    p64.z_obs = torch.zeros((steps, len(gauges)), dtype=torch.float64, device=dev)
    zg, _ = run_forward(p64, t(field(5005)), zbc, keep_checkpoints=False)
    p64.z_obs = zg.clone()
    dm32 = DeviceMesh(h32, device=0)
    prob = problem(dm32)
    prob.z_obs = p64.z_obs
    assert prob.seg == 30
    n0s = torch.stack([t(field(6000 + h)) for h in range(S)]).contiguous()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cals = calibrate_many(prob, n0s, zbc, iters=iters)
    torch.cuda.synchronize()
    print(f"{S} hypotheses x {iters} gradient steps at {host.CELL} cells, {steps} steps: "
          f"{time.perf_counter() - t0:.1f} s", flush=True)

    def misfit64(n):
        zg, _ = run_forward(p64, t(n), zbc, keep_checkpoints=False)
        return float(((zg - p64.z_obs) ** 2).sum())

    first = [misfit64(field(6000 + h)) for h in range(S)]
    final = [misfit64(c.n.cpu().numpy()) for c in cals]
    ratios = np.array(final) / np.array(first)
    print(f"fp64 misfit ratios: median {np.median(ratios):.3f}, max {ratios.max():.3f}", flush=True)
    assert np.all(ratios < 1.0) and np.median(ratios) < 0.5
    old = oracle.set_threads(16)
    try:
        for h in (0, int(np.argsort(ratios)[S // 2])):
            n = cals[h].n.cpu().numpy()
            orc = _oracle_misfit(oracle, host, n, level, gauges, p64.z_obs.cpu().numpy())
            print(f"hypothesis {h}: final misfit GPU fp64 {final[h]:.15e}, oracle {orc:.15e}", flush=True)
            assert abs(final[h] - orc) <= 1e-9 * orc
    finally:
        oracle.set_threads(old)
    dm32.close()
    dm64.close()
