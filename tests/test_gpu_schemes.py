"""GPU: the TILES and CELL work splits, every execution mode and every
boundary kind give identical bits, and the oracle's values.

The 6,675-cell channel (turned by 0.5 rad) gets its wall slots recoded round-robin to KLAS 5
(open), 6 (weir), 7 (breach), 3 (rating curve), 2 (a code with no handler)
and 4 (wall), with weir crests inside the water column and a 4-point rating
curve, so the full boundary handlers run next to KLAS 10 / 1 and the
interior ladder (under variant B breaches fire and turn interior).  After 30 steps:
  * two-kernel, fused TILES (blocks 256 and 128), fused CELL, graph CELL,
    persistent TILES and fused TILES with materialised fluxes: H U V Z W and
    KLAS bit for bit equal, and the materialised FLUX0..3 equal the
    two-kernel path's;
  * against the oracle: <= 1e-12 (fp64) / 1e-6 (fp32) normwise relative.
"""
import numpy as np
import pytest

from conftest import cached_mesh, rare_kinds_mesh, rel_err

pytestmark = pytest.mark.gpu

FIELDS = ("H", "U", "V", "Z", "W")
FLUXES = ("FLUX0", "FLUX1", "FLUX2", "FLUX3")


def _dm(host, variant, **tune):
    import hydro_amd
    from hydro_amd.device import DeviceMesh

    return DeviceMesh(host, variant=variant, device=0, tune=hydro_amd.default_tune(**tune))


@pytest.mark.parametrize("variant", [0, 1], ids=["A", "B"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_schemes_modes_and_rare_kinds(oracle, dtype, variant):
    # A firing KLAS 7 slot divides by the breach depth HB = 0 (the reference
    # reads an uninitialised HB; the Taichi port's 0 is used), an infinite
    # momentum flux that only variant B's CFL step absorbs: breaches fire
    # under variant B and stay closed (crest above the water) under A.
    # The channel is turned by 0.5 rad: with axis-aligned walls a fired
    # breach meets SIDE*COSF == 0, and 0 * inf is a NaN whose sign (which
    # copysign then reads) no arithmetic standard fixes.
    host = rare_kinds_mesh(cached_mesh("channel", dtype=dtype, rotate=0.5), breach_fires=(variant == 1))
    steps = 30
    orc = oracle.OracleRun(host, variant)
    orc.step(steps)
    runs = {}
    cases = (("two_kernel", "two_kernel", {}),
             ("fused-tiles", "fused", dict(scheme=0)),
             ("fused-tiles-128", "fused", dict(scheme=0, block=128)),
             ("fused-cell", "fused", dict(scheme=1)),
             ("graph-cell", "graph", dict(scheme=1, graph_chunk=4)),
             ("persistent-tiles", "persistent", dict(scheme=0)),
             ("fused-tiles-materialized", "fused", dict(scheme=0, materialize_flux=1)))
    for label, mode, tune in cases:
        print(f"case {label} ({np.dtype(dtype).name}, variant {variant})", flush=True)
        dm = _dm(host, variant, **tune)
        dm.step(steps, mode)
        runs[label] = dm.fetch(*FIELDS, "KLAS", *FLUXES)
        dm.close()
    ref = runs["two_kernel"]
    if variant == 1:
        assert np.any(ref["KLAS"] != host.arrays["KLAS"]), "no KLAS 7 slot turned interior"
    for label, r in runs.items():
        for k in FIELDS + ("KLAS",):
            assert np.array_equal(r[k].view(np.uint8), ref[k].view(np.uint8)), f"{label}: {k} differs from two_kernel"
    for k in FLUXES:
        assert np.array_equal(runs["fused-tiles-materialized"][k].view(np.uint8), ref[k].view(np.uint8)), k
    tol = 1e-12 if dtype == np.float64 else 1e-6
    for k in FIELDS:
        err = rel_err(ref[k], orc[k])
        print(f"{k}: rel {err:.3e} vs oracle")
        assert err <= tol, k
    assert np.array_equal(ref["KLAS"], orc["KLAS"])
