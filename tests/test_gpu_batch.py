"""GPU: batched ensembles give exactly the single-mesh results.

Members of uneven size (2K-20K cells, own bed, Manning and hydrograph) are
advanced together with one launch per step, a replayed graph, or a single
persistent launch; every member's state must equal, bit for bit, the state
of the same mesh stepped alone (same arithmetic per cell).  Members may start
at different times.  The fp32 members are also compared with their own fp64
runs (the reference's fp32-vs-fp64 envelope).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_MEMBERS = {}


def _members(dtype):
    import hydro_amd.mesh as M

    key = np.dtype(dtype).str
    if key not in _MEMBERS:
        _MEMBERS[key] = M.make_ensemble(seed=77, n=6, lo=2000, hi=20000, steps=120, dtype=dtype, workers=1)
    return _MEMBERS[key]


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("mode", ["fused", "graph", "persistent"])
def test_batch_equals_single_mesh_runs(dtype, mode):
    import hydro_amd
    from hydro_amd.device import DeviceMesh
    from hydro_amd.ensemble import Batch

    hosts = _members(dtype)
    tune = hydro_amd.default_tune(block=128, graph_chunk=8)
    singles = []
    for i, h in enumerate(hosts):
        dm = DeviceMesh(h, device=0, tune=tune)
        if i % 2:
            dm.step(3, "fused")            # members at different times
        dm.step(50, "fused")
        singles.append(dm.fetch("H", "U", "V", "Z"))
        dm.close()
    dms = [DeviceMesh(h, device=0, tune=tune) for h in hosts]
    for i, dm in enumerate(dms):
        if i % 2:
            dm.step(3, "fused")
    b = Batch(dms, tune=tune)
    b.step(50, mode)
    b.sync()
    for i, dm in enumerate(dms):
        got = dm.fetch("H", "U", "V", "Z")
        for k in ("H", "U", "V", "Z"):
            assert np.array_equal(got[k], singles[i][k]), f"member {i} ({hosts[i].CELL} cells) {k} differs"
        assert dm.get_time()[1] == 1 + (50 + (3 if i % 2 else 0)) % (hosts[i].steps_per_day - 1)
    b.close()
    for dm in dms:
        dm.close()


def test_batch_fp32_members_within_fp64_envelope():
    from hydro_amd.device import DeviceMesh
    from hydro_amd.ensemble import Batch

    h32, h64 = _members(np.float32), _members(np.float64)
    dms = [DeviceMesh(h, device=0) for h in h32]
    b = Batch(dms)
    b.step(100, "persistent")
    b.sync()
    worst = 0.0
    for i, (dm, h) in enumerate(zip(dms, h64)):
        ref = DeviceMesh(h, device=0)
        ref.step(100, "fused")
        H64 = ref.fetch("H")["H"]
        H32 = dm.fetch("H")["H"].astype(np.float64)
        worst = max(worst, float(np.max(np.abs(H32 - H64))))
        ref.close()
    print(f"fp32 batch vs fp64 single runs, 100 steps: max |dH| = {worst:.3e} m")
    assert worst < 1e-3
    b.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_small_tiles_under_wide_workgroups(dtype):
    """Meshes with fewer full tiles than CUs get tiles of fewer cells than
    the workgroup has threads (here 64-79 cells under 256 threads): the
    fused and persistent steps must still give the two-kernel bits (every
    thread of the last tile loads task entries past its tile, which the
    task arrays must hold)."""
    import hydro_amd
    from hydro_amd.device import DeviceMesh

    tune = hydro_amd.default_tune(block=256)
    for h in _members(dtype):
        res = {}
        for mode in ("two_kernel", "fused", "persistent"):
            dm = DeviceMesh(h, device=0, tune=tune)
            dm.step(50, mode)
            res[mode] = dm.fetch("H", "U", "V", "Z")
            dm.close()
        for mode in ("fused", "persistent"):
            for k in ("H", "U", "V", "Z"):
                assert np.array_equal(res[mode][k], res["two_kernel"][k]), f"{h.CELL} cells {mode} {k}"
