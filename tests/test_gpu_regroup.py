"""GPU: regrouping the task plan by the current wet/dry state (hydro_regroup).

hydro_prepare orders each tile's pair tasks by the wet/dry branches their
two sides take in the state it finds; hydro_regroup does the same on the
device for the current state, permuting each tile's pair tasks among the
entries that hold pair tasks.

* A state whose wet/dry pattern differs from the one the plan was made for
  leaves many pair chunks mixing classes; after hydro_regroup at most two
  per tile do (the class boundaries), and the chunk count is unchanged.
* Which entry runs a pair changes no result: steps after a regroup equal,
  bit for bit, the same steps on a mesh that was not regrouped, in the
  persistent and fused modes, fp64 and fp32, and equal the two-kernel path
  (which does not use the plan).
"""
import numpy as np
import pytest

from conftest import cached_mesh

pytestmark = pytest.mark.gpu

FIELDS = ("H", "U", "V", "Z")


def _tune():
    import hydro_amd

    return hydro_amd.default_tune(block=1024, tile_cells=-1, resident_launch=1)


def _reshuffled_state(host, seed):
    """The mesh's state with a different wet/dry pattern: a random third of
    the cells dry (H = HM1 / 2 at rest), the rest as they were."""
    st = {k: v.copy() for k, v in host.state().items()}
    rng = np.random.default_rng(seed)
    dry = rng.random(host.CELL) < 1 / 3
    dt = st["H"].dtype
    st["H"][dry] = dt.type(host.HM1 * 0.5)
    st["U"][dry] = 0
    st["V"][dry] = 0
    st["Z"][dry] = (host.ZBC[dry] + st["H"][dry]).astype(dt)
    return st


def _run(host, state, steps, mode, regroup_at=()):
    import hydro_amd
    from hydro_amd.device import DeviceMesh

    dm = DeviceMesh(host, device=0, tune=_tune())
    dm.load_state(state)
    dm.set_time(0, 1)
    done = 0
    for at in list(regroup_at) + [steps]:
        if at > done:
            dm.step(at - done, mode)
            done = at
        if at < steps:
            dm.regroup()
    out = dm.fetch(*FIELDS)
    dm.close()
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_regroup_groups_the_current_state(dtype):
    from hydro_amd.device import DeviceMesh

    host = cached_mesh("floodplain_207234", dtype=dtype)
    dm = DeviceMesh(host, device=0, tune=_tune())
    tiles = dm.plan_stats()[0]
    pc0, mixed0 = dm.plan_grouping()
    assert pc0 > 0 and mixed0 <= 2 * tiles          # the planner's own grouping
    dm.load_state(_reshuffled_state(host, 7))
    dm.set_time(0, 1)
    pc1, mixed1 = dm.plan_grouping()
    assert pc1 == pc0 and mixed1 > 4 * tiles         # the new pattern mixes most chunks
    dm.regroup()
    pc2, mixed2 = dm.plan_grouping()
    print(f"tiles {tiles}: pair chunks {pc0}, mixed {mixed0} (planned), {mixed1} (new state), {mixed2} (regrouped)")
    assert pc2 == pc0 and mixed2 <= 2 * tiles
    dm.close()


@pytest.mark.parametrize("mode", ["persistent", "fused"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_regroup_changes_no_result(dtype, mode):
    host = cached_mesh("floodplain_207234", dtype=dtype)
    state = _reshuffled_state(host, 11)
    got = _run(host, state, 60, mode, regroup_at=(0, 25, 40))
    want = _run(host, state, 60, mode)
    for k in FIELDS:
        assert np.array_equal(got[k], want[k]), k
    ref = _run(host, state, 60, "two_kernel")
    for k in FIELDS:
        assert np.array_equal(got[k], ref[k]), k


SMALL = ("one cell", "three cells, triangles", "65 cells", "257-cell strip", "triangles and quads", "dry start")


@pytest.mark.parametrize("name", SMALL)
def test_frequent_regroups_on_small_meshes(name):
    """Regrouping every 3 steps (hydro_tune.regroup_steps) on the small and
    degenerate meshes: every reordered mode equals the two-kernel path bit
    for bit after 30 steps (across a day boundary)."""
    import hydro_amd
    from hydro_amd.device import DeviceMesh
    from test_gpu_edge_cases import CASES, grid_mesh

    host = grid_mesh(**dict(CASES)[name])
    ref = DeviceMesh(host, device=0)
    ref.step(30, "two_kernel")
    want = ref.fetch(*FIELDS)
    ref.close()
    for mode in ("fused", "graph", "persistent"):
        dm = DeviceMesh(host, device=0, tune=hydro_amd.default_tune(regroup_steps=3))
        dm.step(30, mode)
        got = dm.fetch(*FIELDS)
        dm.close()
        for k in FIELDS:
            assert np.array_equal(got[k], want[k]), (mode, k)


def test_frequent_regroups_with_every_boundary_kind():
    """The same on a mesh with every boundary kind (the full kernel variant),
    breaches opening mid-run."""
    import hydro_amd
    from conftest import rare_kinds_mesh
    from hydro_amd.device import DeviceMesh

    host = rare_kinds_mesh(cached_mesh("channel_6675"), breach_fires=True)
    ref = DeviceMesh(host, device=0)
    ref.step(40, "two_kernel")
    want = ref.fetch(*FIELDS)
    ref.close()
    for mode in ("fused", "persistent"):
        dm = DeviceMesh(host, device=0, tune=hydro_amd.default_tune(regroup_steps=4))
        dm.step(40, mode)
        got = dm.fetch(*FIELDS)
        dm.close()
        for k in FIELDS:
            assert np.array_equal(got[k], want[k]), (mode, k)


@pytest.mark.parametrize("mode", ["fused", "graph", "persistent"])
def test_batch_regroups_change_no_result(mode):
    """A batch regrouping every member every 4 steps equals each member's
    two-kernel run bit for bit after 30 steps."""
    import hydro_amd
    import hydro_amd.mesh as M
    from hydro_amd.device import DeviceMesh
    from hydro_amd.ensemble import Batch

    hosts = M.make_ensemble(seed=91, n=4, lo=3000, hi=9000, steps=80, dtype=np.float64, workers=1)
    tune = hydro_amd.default_tune(block=128)
    dms = [DeviceMesh(h, device=0, tune=tune) for h in hosts]
    b = Batch(dms, tune=hydro_amd.default_tune(block=128, regroup_steps=4))
    b.step(30, mode)
    b.sync()
    got = [dm.fetch(*FIELDS) for dm in dms]
    b.close()
    for dm in dms:
        dm.close()
    for h, g in zip(hosts, got):
        ref = DeviceMesh(h, device=0)
        ref.step(30, "two_kernel")
        want = ref.fetch(*FIELDS)
        ref.close()
        for k in FIELDS:
            assert np.array_equal(g[k], want[k]), k
