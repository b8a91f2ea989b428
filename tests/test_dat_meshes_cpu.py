"""CPU: the reference's own meshes read from its .DAT directories
(hydro_amd.dat), F1_text (6,675 cells; discharge edges and an unknown KLAS
code 13) and F2_24K_text (24,020 cells; 169 water-level edges with their
series), run through the device source compiled for the host against the
oracle: geometry equal to the oracle's preCalculate bit for bit, then 40 steps
of fluxes and state within 1e-12 normwise (fp64).  Needs the reference
checkout (the data directories); skipped without it."""
import ctypes
import os

import numpy as np
import pytest

from test_host_step_cpu import _desc, _diff, host_step  # noqa: F401  (fixture)

DATA = "/root/reference/data"

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(DATA, "F1_text")),
                                reason="reference data directories not present")

_MESHES = {}


def _mesh(name):
    from hydro_amd.dat import load_dat

    if name not in _MESHES:
        _MESHES[name] = load_dat(os.path.join(DATA, name), ndays=2)
    return _MESHES[name]


@pytest.mark.parametrize("name", ["F1_text", "F2_24K_text"])
def test_geometry_equals_oracle_precalculate(oracle, name):
    host = _mesh(name)
    a = host.arrays
    # the reader's inputs again, through the oracle's restatement of preCalculate
    from hydro_amd.dat import _table

    root = os.path.join(DATA, name)
    NOD = len(a["XP"])
    xy = _table(root, "PXY", NOD, 3)
    z0 = _table(root, "INITIALLEVEL", host.CELL, 1)[:, 0]
    n = _table(root, "CV", host.CELL, 1)[:, 0]
    ref = oracle.precalculate(host.CELL, NOD, a["NAP"], xy[:, 1], xy[:, 2], a["ZBC"], z0.copy(), n, host.HM1)
    for k in ("SIDE", "SINF", "COSF", "SLCOS", "SLSIN", "AREA", "ZB1", "H", "FNC", "Z"):
        assert np.array_equal(a[k], ref[k]), k
    assert np.all(a["AREA"] > 0)


@pytest.mark.parametrize("name", ["F1_text", "F2_24K_text"])
def test_device_source_on_reference_meshes(host_step, oracle, name):  # noqa: F811
    host = _mesh(name)
    arrays = {k: v.copy() for k, v in host.arrays.items()}
    d = _desc(host, arrays, 0)
    orc = oracle.OracleRun(host, 0)
    spd = host.steps_per_day
    day, kt = 0, 1
    worst = 0.0
    for step in range(1, 41):
        host_step.host_calculate_flux(ctypes.byref(d), kt, day)
        orc.calculate_flux(kt, day)
        for k in ("FLUX0", "FLUX1", "FLUX2", "FLUX3"):
            e = _diff(arrays[k], orc[k])
            worst = max(worst, e)
            assert e <= 1e-12, f"step {step} {k}: {e:.3e}"
        host_step.host_update_cell(ctypes.byref(d))
        orc.update_cell()
        for k in ("H", "U", "V", "Z", "W"):
            e = _diff(arrays[k], orc[k])
            worst = max(worst, e)
            assert e <= 1e-12, f"step {step} {k}: {e:.3e}"
        kt += 1
        if kt >= spd:
            kt, day = 1, day + 1
    moved = float(np.abs(arrays["U"]).max())
    print(f"{name}: worst normwise difference over 40 steps {worst:.2e}, max |U| {moved:.3e}")
    assert moved > 0
