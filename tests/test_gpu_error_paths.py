"""GPU: error reporting of the persistent launches, and refreshed statics.

* A persistent launch whose waits between steps exceed their spin limit
  leaves early; the next call that reads the state must report it (single
  mesh and batch), and the mesh must step normally afterwards.  The
  condition is forced from the host with a spin limit of one sleep: every
  wait still ends, so every kernel exits normally.
* After the caller changes AREA (and ZBC, FNC) and calls
  hydro_refresh_statics, the fused step (which reads DT / AREA prepared per
  cell) must equal the two-kernel step (which forms it from AREA) bit for
  bit.
"""
import numpy as np
import pytest

from conftest import cached_mesh

pytestmark = pytest.mark.gpu


def _state_equal(a, b):
    return all(np.array_equal(a[k], b[k]) for k in ("H", "U", "V", "Z"))


def test_persistent_spin_limit_is_reported_single_mesh():
    import hydro_amd
    from hydro_amd import _lib
    from hydro_amd.device import DeviceMesh

    host = cached_mesh("floodplain_24020")
    init = host.state()
    tight = hydro_amd.default_tune(scheme=_lib.SCHEME_CELL, spin_limit=1)
    dm = DeviceMesh(host, device=0, tune=tight)
    dm.step(30, "persistent")
    with pytest.raises(_lib.HydroError) as e:
        dm.sync()
    assert e.value.status == _lib.ERR_HIP and "spin limit" in str(e.value)
    # the flag was cleared: with the default limit the mesh steps normally
    dm.prepare(hydro_amd.default_tune(scheme=_lib.SCHEME_CELL))
    dm.load_state(init)
    dm.set_time(0, 1)
    dm.step(20, "persistent")
    got = dm.fetch("H", "U", "V", "Z")
    dm.close()
    ref = DeviceMesh(host, device=0)
    ref.step(20, "fused")
    want = ref.fetch("H", "U", "V", "Z")
    ref.close()
    assert _state_equal(got, want)


def test_persistent_spin_limit_is_reported_batch():
    import hydro_amd
    import hydro_amd.mesh as M
    from hydro_amd import _lib
    from hydro_amd.device import DeviceMesh
    from hydro_amd.ensemble import Batch

    hosts = M.make_ensemble(seed=91, n=3, lo=3000, hi=9000, steps=80, dtype=np.float64, workers=1)
    tune = hydro_amd.default_tune(block=128)
    dms = [DeviceMesh(h, device=0, tune=tune) for h in hosts]
    b = Batch(dms, tune=hydro_amd.default_tune(block=128, spin_limit=1))
    b.step(30, "persistent")
    with pytest.raises(_lib.HydroError) as e:
        b.sync()
    assert e.value.status == _lib.ERR_HIP and "spin limit" in str(e.value)
    b.close()
    for dm, h in zip(dms, hosts):
        dm.load_state(h.state())
        dm.set_time(0, 1)
    b = Batch(dms, tune=tune)
    b.step(20, "persistent")
    b.sync()
    got = [dm.fetch("H", "U", "V", "Z") for dm in dms]
    b.close()
    for dm in dms:
        dm.close()
    for h, g in zip(hosts, got):
        ref = DeviceMesh(h, device=0, tune=tune)
        ref.step(20, "fused")
        assert _state_equal(g, ref.fetch("H", "U", "V", "Z"))
        ref.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_refresh_statics_rebuilds_dt_over_area(dtype):
    import torch

    from hydro_amd import _lib
    from hydro_amd.device import DeviceMesh

    host = cached_mesh("channel_6675", dtype=dtype)
    init = host.state()
    rng = np.random.default_rng(5)
    area = (host.AREA * rng.uniform(0.7, 1.3, host.CELL)).astype(dtype)
    fnc = (host.FNC * rng.uniform(0.8, 1.2, host.CELL)).astype(dtype)
    out = {}
    for mode in ("fused", "two_kernel"):
        dm = DeviceMesh(host, device=0)
        dm.step(5, mode)                 # prepared layout in use before the change
        dm.load_state(init)
        dm.set_time(0, 1)
        dm.t["AREA"].copy_(torch.from_numpy(area))
        dm.t["FNC"].copy_(torch.from_numpy(fnc))
        _lib.check(dm.lib.hydro_refresh_statics(dm.h, dm.stream()))
        dm.step(25, mode)
        out[mode] = dm.fetch("H", "U", "V", "Z")
        dm.close()
    assert _state_equal(out["fused"], out["two_kernel"])
    # and the change took effect
    dm = DeviceMesh(host, device=0)
    dm.step(25, "fused")
    assert not np.array_equal(dm.fetch("H")["H"], out["fused"]["H"])
    dm.close()
