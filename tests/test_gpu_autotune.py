"""GPU: the autotuners pick a tuning without changing the answer.

hydro_autotune (one mesh) and ensemble.autotune_batch (a batch) time their
candidates from the caller's state and must leave that state and the step
time untouched; every candidate is an exact re-arrangement of the same
arithmetic, so stepping on with the chosen tuning gives the two-kernel bits.
"""
import numpy as np
import pytest

from conftest import cached_mesh

pytestmark = pytest.mark.gpu


def test_mesh_autotune_keeps_state_and_bits():
    import hydro_amd
    from hydro_amd.device import DeviceMesh

    host = cached_mesh("floodplain_24020")
    dm = DeviceMesh(host, device=0)
    dm.step(7, "fused")
    before = dm.fetch("H", "U", "V", "Z")
    t_before = dm.get_time()
    best, table = dm.autotune("fused", steps=20)
    for tune, us in table:
        print(f"block {tune.block} tile {tune.tile_cells} scheme {tune.scheme} wps {tune.waves_per_simd}: {us:.2f} us")
    assert all(us > 0 for _, us in table)
    assert (best.block, best.tile_cells, best.scheme, best.waves_per_simd) in [
        (t.block, t.tile_cells, t.scheme, t.waves_per_simd) for t, us in table if us == min(u for _, u in table)]
    assert dm.get_time() == t_before
    after = dm.fetch("H", "U", "V", "Z")
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    dm.step(30, "fused")
    got = dm.fetch("H", "U", "V", "Z")
    dm.close()
    ref = DeviceMesh(host, device=0)
    ref.step(37, "two_kernel")
    want = ref.fetch("H", "U", "V", "Z")
    ref.close()
    for k in want:
        assert np.array_equal(got[k], want[k]), k


def test_batch_autotune_keeps_states():
    import hydro_amd.mesh as M
    from hydro_amd.device import DeviceMesh
    from hydro_amd.ensemble import Batch, autotune_batch

    hosts = M.make_ensemble(seed=77, n=6, lo=2000, hi=20000, steps=120, dtype=np.float32, workers=1)
    dms = [DeviceMesh(h, device=0) for h in hosts]
    dms[1].step(3, "fused")
    before = [dm.fetch("H", "U", "V", "Z") for dm in dms]
    times = [dm.get_time() for dm in dms]
    (block, mode), table = autotune_batch(dms, steps=10)
    print(table)
    assert (block, mode) in table and all(v > 0 for v in table.values())
    for dm, st, t in zip(dms, before, times):
        assert dm.get_time() == t
        got = dm.fetch("H", "U", "V", "Z")
        for k in st:
            assert np.array_equal(got[k], st[k]), k
    b = Batch(dms, hydro_amd_tune(block))
    b.step(20, mode)
    b.sync()
    b.close()
    for dm in dms:
        dm.close()


def hydro_amd_tune(block):
    import hydro_amd

    return hydro_amd.default_tune(block=block)
