"""GPU: the scenario loops (several Manning fields of one mesh per launch)
against the one-field loops, and the batched calibration against the
one-by-one calibration.

On the 24,020-cell floodplain (60 steps, checkpoints every 8, 16 gauges,
three Manning fields), in fp64 and fp32: every scenario's gauge levels and
both gradients equal, bit for bit, those of the one-field forward run and
adjoint sweep with that field, and calibrate_many gives each hypothesis the
misfit sequence and final field of calibrate on it alone.
"""
import numpy as np
import pytest

from conftest import cached_mesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_scenario_loops_equal_one_field_loops(dtype):
    import torch

    from hydro_amd.adjoint import MisfitProblem, run_backward, run_backward_scen, run_forward, run_forward_scen
    from hydro_amd.calibrate import calibrate, calibrate_many, smooth_field
    from hydro_amd.device import DeviceMesh

    h64 = cached_mesh("floodplain_24020")
    host = cached_mesh("floodplain_24020", dtype=dtype)
    dm = DeviceMesh(host, device=0)
    dev = dm.device
    f64 = lambda x: torch.tensor(np.asarray(x, dtype=np.float64), dtype=torch.float64, device=dev)
    rng = np.random.default_rng(7007)
    level = h64.Z.copy()
    level[h64.H <= h64.HM1] = h64.ZBC[h64.H <= h64.HM1]
    gauges = rng.choice(np.nonzero(h64.H > 0.5)[0], size=16, replace=False).astype(np.int32)
    steps = 60
    prob = MisfitProblem(dm=dm, level0=f64(level), U0=f64(h64.U), V0=f64(h64.V),
                         gauges=torch.tensor(gauges, dtype=torch.int32, device=dev), steps=steps, segment=8)
    zbc = f64(h64.ZBC)
    field = lambda seed: np.clip(smooth_field(h64.xc, h64.yc, seed, 0.04, 0.01, 800.0), 0.005, 0.2)
    prob.z_obs = torch.zeros((steps, 16), dtype=torch.float64, device=dev)
    zg, _ = run_forward(prob, f64(field(1)), zbc, keep_checkpoints=False)
    prob.z_obs = zg.clone()
    ns = torch.stack([f64(field(10 + k)) for k in range(3)]).contiguous()

    zs, buf = run_forward_scen(prob, ns, zbc)
    gs_n, gs_z = run_backward_scen(prob, buf, zs, ns, zbc)
    del buf
    for k in range(3):
        z1, b1 = run_forward(prob, ns[k].clone(), zbc)
        g1_n, g1_z = run_backward(prob, b1, z1, ns[k].clone(), zbc)
        assert torch.equal(zs[:, k, :], z1), f"scenario {k} gauge levels"
        assert torch.equal(gs_n[k], g1_n), f"scenario {k} Manning gradient"
        assert torch.equal(gs_z[k], g1_z), f"scenario {k} bed gradient"
        assert float(g1_n.abs().max()) > 0

    cals = calibrate_many(prob, ns, zbc, iters=3)
    for k in range(3):
        one = calibrate(prob, ns[k].clone(), zbc, iters=3)
        print(f"hypothesis {k}: {[f'{x:.6e}' for x in one.losses]}")
        assert cals[k].losses == one.losses and torch.equal(cals[k].n, one.n)
        assert one.best < one.losses[0]
    dm.close()
