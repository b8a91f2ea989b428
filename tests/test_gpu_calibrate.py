"""GPU: Manning calibration by gradient descent (hydro_amd.calibrate) lowers
the gauge misfit, and the misfit the GPU reports for the calibrated field
equals the fp64 CPU oracle's for the same field (24,020-cell floodplain,
200 steps, checkpoints every 15, 32 gauges, two starting hypotheses, three
gradient steps each)."""
import numpy as np
import pytest

from conftest import cached_mesh, rel_err

pytestmark = pytest.mark.gpu


def _oracle_misfit(oracle, host, n, level, gauges, z_obs):
    h = host.copy()
    a = h.arrays
    a["FNC"][:] = 9.81 * n * n
    a["ZB1"][:] = a["ZBC"] + host.HM1
    wet = level > a["ZBC"]
    a["H"][:] = np.where(wet, level - a["ZBC"], host.HM1)
    a["Z"][:] = np.where(wet, level, a["ZBC"] + host.HM1)
    orc = oracle.OracleRun(h)
    zg = np.empty_like(z_obs)
    for s in range(z_obs.shape[0]):
        orc.step(1)
        zg[s] = orc["Z"][gauges]
    return float(((zg - z_obs) ** 2).sum())


def test_calibration_lowers_misfit_and_matches_oracle(oracle):
    import torch

    from hydro_amd.adjoint import MisfitProblem, run_forward
    from hydro_amd.calibrate import calibrate, smooth_field
    from hydro_amd.device import DeviceMesh

    host = cached_mesh("floodplain_24020")
    steps = 200
    dm = DeviceMesh(host, device=0)
    dev = dm.device
    t = lambda x, dt=torch.float64: torch.tensor(x, dtype=dt, device=dev)
    rng = np.random.default_rng(4004)
    level = host.Z.copy()
    level[host.H <= host.HM1] = host.ZBC[host.H <= host.HM1]
    gauges = rng.choice(np.nonzero(host.H > 0.5)[0], size=32, replace=False).astype(np.int32)
    prob = MisfitProblem(dm=dm, level0=t(level), U0=t(host.U), V0=t(host.V), gauges=t(gauges, torch.int32),
                         steps=steps, segment=15)
    zbc = t(host.ZBC.copy())
    truth = np.clip(smooth_field(host.xc, host.yc, 5005, 0.04, 0.01, 800.0), 0.005, 0.2)
    prob.z_obs = torch.zeros((steps, 32), dtype=torch.float64, device=dev)
    zg, _ = run_forward(prob, t(truth), zbc, keep_checkpoints=False)
    prob.z_obs = zg.clone()
    for hyp in range(2):
        n0 = np.clip(smooth_field(host.xc, host.yc, 6000 + hyp, 0.04, 0.01, 800.0), 0.005, 0.2)
        cal = calibrate(prob, t(n0), zbc, iters=3)
        print(f"hypothesis {hyp}: misfit {[f'{x:.4e}' for x in cal.losses]}")
        assert cal.best < 0.5 * cal.losses[0]
        n_final = cal.n.cpu().numpy()
        orc = _oracle_misfit(oracle, host, n_final, level, gauges, prob.z_obs.cpu().numpy())
        print(f"  final misfit GPU {cal.best:.15e} oracle {orc:.15e}")
        assert abs(cal.best - orc) <= 1e-9 * orc
    dm.close()
