"""GPU: every cell order of the internal layout (NONE, RCM, HILBERT, TILES)
under both work splits gives the two-kernel step's bits (which equal the CPU
oracle to the step's parity bar), because the primary side of an edge is
decided on the caller's cell ids whatever the order; the autotuner searches
the cell order too."""
import numpy as np
import pytest

from conftest import cached_mesh, rare_kinds_mesh, rel_err

pytestmark = pytest.mark.gpu

ORDERS = {"none": 0, "rcm": 1, "hilbert": 2, "tiles": 3}


def _run(host, mode, tune, steps, variant=0):
    from hydro_amd.device import DeviceMesh

    dm = DeviceMesh(host, variant=variant, device=0, tune=tune)
    dm.step(steps, mode)
    out = dm.fetch("H", "U", "V", "Z", "KLAS")
    stats = dm.plan_stats() if mode != "two_kernel" else None
    dm.close()
    return out, stats


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("mesh", ["channel", "rare_kinds"])
def test_every_cell_order_gives_the_reference_bits(dtype, mesh, oracle):
    import hydro_amd
    from hydro_amd import _lib

    variant = 0
    host = cached_mesh("channel_6675", dtype=dtype)
    if mesh == "rare_kinds":
        # every boundary kind, breaches opening under variant B, whose CFL
        # step absorbs their infinite momentum flux; the channel turned so
        # no wall is axis-aligned (as in test_gpu_schemes.py)
        host = rare_kinds_mesh(cached_mesh("channel", dtype=dtype, rotate=0.5), breach_fires=True)
        variant = 1
    steps = 30
    ref, _ = _run(host, "two_kernel", hydro_amd.default_tune(), steps, variant)
    r = oracle.OracleRun(host, variant)
    r.step(steps)
    bar = 1e-12 if dtype == np.float64 else 1e-6
    assert max(rel_err(ref[k], r[k]) for k in "HUVZ") <= bar
    for name, order in ORDERS.items():
        for scheme in (_lib.SCHEME_TILES, _lib.SCHEME_CELL):
            tune = hydro_amd.default_tune(reorder=order, scheme=scheme)
            for mode in ("fused", "persistent"):
                got, stats = _run(host, mode, tune, steps, variant)
                for k in ("H", "U", "V", "Z", "KLAS"):
                    assert np.array_equal(got[k].view(np.uint8), ref[k].view(np.uint8)), (name, scheme, mode, k)
            if scheme == _lib.SCHEME_TILES:
                print(f"{mesh} {name}: {stats[0]} tiles, {stats[1]} pair tasks, {stats[2]} one-slot tasks, "
                      f"{stats[3]:.2f} edges per thread", flush=True)


def test_order_quality_and_autotune_axis():
    import hydro_amd
    from hydro_amd import _lib
    from hydro_amd.device import DeviceMesh

    host = cached_mesh("floodplain_24020")
    pairs = {}
    for name, order in ORDERS.items():
        dm = DeviceMesh(host, device=0, tune=hydro_amd.default_tune(reorder=order, tile_cells=256))
        pairs[name] = dm.plan_stats()[1]
        dm.close()
    print("pair tasks (edges solved once inside a tile):", pairs, flush=True)
    # compact orders keep most edges inside a tile; the caller's order does not
    assert pairs["tiles"] > pairs["hilbert"] > pairs["none"]
    assert pairs["rcm"] > pairs["none"]
    dm = DeviceMesh(host, device=0)
    st0 = dm.fetch("H", "U", "V", "Z")
    best, table = dm.autotune("fused", steps=20)
    orders = {t.reorder for t, us in table if us > 0}
    assert {_lib.REORDER_TILES, _lib.REORDER_RCM, _lib.REORDER_HILBERT} <= orders
    print("autotune:", [(t.reorder, t.scheme, t.block, t.tile_cells, round(us, 2)) for t, us in table], flush=True)
    st1 = dm.fetch("H", "U", "V", "Z")
    assert all(np.array_equal(st0[k], st1[k]) for k in "HUVZ")     # state left as it was
    dm.close()


def test_hilbert_needs_coordinates():
    import hydro_amd
    from hydro_amd import _lib
    from hydro_amd.device import DeviceMesh

    host = cached_mesh("channel_6675").copy()
    host.xc = host.yc = None
    dm = DeviceMesh(host, device=0)
    with pytest.raises(_lib.HydroError) as e:
        dm.prepare(hydro_amd.default_tune(reorder=_lib.REORDER_HILBERT))
    assert e.value.status == _lib.ERR_UNSUPPORTED
    dm.close()
