"""GPU: device-reduced diagnostics (hydro_diagnostics, DPP wavefront sums).

* Every sum equals the correctly rounded sum (math.fsum) of the reference
  checker's own terms (scripts/check_correctness.py:620-655), formed on the
  host from the device's state and fluxes: fp64 and fp32 meshes, two-kernel
  and fused steps, and on the 207,234-cell floodplain (203 workgroup records).
* The same figures of the oracle's state after the same steps agree to the
  step's own parity bar (1e-12 relative in fp64).
* Non-finite and infinite depths are counted, and the depth range skips them.
* Mass budget: on the channel with its outlet closed and every cell wet, the
  volume change of a step is DT times the KLAS 10 inflow the device reduces
  from that step's fluxes.
"""
import math

import numpy as np
import pytest

from conftest import cached_mesh, rel_err

pytestmark = pytest.mark.gpu

G = 9.81


def checker_terms(H, U, V, area, F0, side, klas):
    """The checker's terms, in its order of operations, in double."""
    H, U, V, a = (np.asarray(x, dtype=np.float64) for x in (H, U, V, area))
    F0, side = np.asarray(F0, np.float64), np.asarray(side, np.float64)
    return {
        "mass": H * a,
        "momentum_x": H * U * a,
        "momentum_y": H * V * a,
        "kinetic_energy": 0.5 * H * (U * U + V * V) * a,
        "potential_energy": 0.5 * G * H * H * a,
        "klas10_inflow": (F0 * side)[klas == 10],
        "klas1_inflow": (F0 * side)[klas == 1],
    }


def expected(host, st):
    t = checker_terms(st["H"], st["U"], st["V"], host.AREA, st["FLUX0"], host.SIDE, host.KLAS)
    out = {k: math.fsum(v.tolist()) for k, v in t.items()}
    H = np.asarray(st["H"], np.float64)
    fin = np.isfinite(H)
    out["nonfinite_h"] = float(np.count_nonzero(~fin))
    out["inf_h"] = float(np.count_nonzero(np.isinf(H)))
    out["h_min"] = float(H[fin].min())
    out["h_max"] = float(H[fin].max())
    return out, t


def assert_sums_exact(got, want, terms):
    for k, v in want.items():
        if k in terms:
            # correctly rounded: equal, or (near-total cancellation) within a
            # hair of the terms' magnitude
            tol = 1e-25 * float(np.sum(np.abs(terms[k]))) if terms[k].size else 0.0
            assert got[k] == v or abs(got[k] - v) <= tol, (k, got[k], v)
        else:
            assert got[k] == v, (k, got[k], v)


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("mode", ["two_kernel", "fused"])
def test_diagnostics_equal_fsum_of_checker_terms(dtype, mode, oracle):
    import hydro_amd
    from hydro_amd.device import DeviceMesh

    host = cached_mesh("channel_6675", dtype=dtype)
    dm = DeviceMesh(host, device=0, tune=hydro_amd.default_tune(materialize_flux=1))
    dm.step(7, mode)
    got = dm.diagnostics()
    st = dm.fetch("H", "U", "V", "Z", "FLUX0")
    dm.close()
    want, terms = expected(host, st)
    assert_sums_exact(got, want, terms)
    assert got["klas10_inflow"] < 0 and got["nonfinite_h"] == 0
    # the oracle's state after the same steps, to the step's parity bar
    r = oracle.OracleRun(host)
    r.step(7)
    ow, _ = expected(host, {k: r[k] for k in ("H", "U", "V", "FLUX0")})
    bar = 1e-12 if dtype == np.float64 else 1e-5
    for k in ("mass", "potential_energy", "klas10_inflow", "klas1_inflow", "h_min", "h_max"):
        assert abs(got[k] - ow[k]) <= bar * max(abs(ow[k]), 1e-300), (k, got[k], ow[k])


def test_diagnostics_at_scale_207k():
    from hydro_amd.device import DeviceMesh

    host = cached_mesh("floodplain_207234")
    dm = DeviceMesh(host, device=0)
    dm.step(20, "two_kernel")
    got = dm.diagnostics()
    st = dm.fetch("H", "U", "V", "Z", "FLUX0")
    dm.close()
    want, terms = expected(host, st)
    assert_sums_exact(got, want, terms)


def test_diagnostics_count_nonfinite_depths():
    from hydro_amd.device import DeviceMesh

    host = cached_mesh("channel_6675")
    st = host.state()
    H = st["H"].copy()
    H[[5, 77, 4000]] = np.nan
    H[[9, 6000]] = np.inf
    H[[10]] = -np.inf
    H[123] = 1e-6                      # the smallest finite depth
    H[321] = 55.0                      # the largest
    st["H"] = H
    dm = DeviceMesh(host, device=0)
    dm.load_state(st)
    got = dm.diagnostics()
    dm.close()
    assert got["nonfinite_h"] == 6 and got["inf_h"] == 3
    assert got["h_min"] == 1e-6 and got["h_max"] == 55.0
    assert math.isnan(got["mass"]) or math.isinf(got["mass"])


def test_mass_budget_from_device_inflow():
    """Closed outlet, all cells wet: V(n+1) - V(n) = -DT * (KLAS 10 inflow of step n+1's fluxes)."""
    from hydro_amd.device import DeviceMesh

    host = cached_mesh("channel_6675").copy()
    k = host.arrays["KLAS"]
    k[k == 1] = 4
    assert host.H.min() > 0.5
    dm = DeviceMesh(host, device=0)
    v_prev = dm.diagnostics()["mass"]
    checked = 0
    for step in range(12):
        dm.step(1, "two_kernel")
        d = dm.diagnostics()
        q = -d["klas10_inflow"]            # into the domain, m^3/s
        assert q > 10.0
        assert abs((d["mass"] - v_prev) / host.DT - q) <= 1e-9 * q, (step, d["mass"] - v_prev, q)
        v_prev = d["mass"]
        checked += 1
    dm.close()
    assert checked == 12
