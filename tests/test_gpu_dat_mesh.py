"""GPU: the reference's own F1_text mesh (6,675 cells, read from its .DAT
directory by hydro_amd.dat; inputs kept in tests/golden/f1_text_mesh.npz by
tests/golden/make_dat_fixture.py), 200 fp64 steps in the two-kernel, fused
and persistent modes: every mode gives the same bits, within 1e-12
normwise of the CPU oracle.  The mesh has discharge edges and KLAS code 13,
which has no handler (FLUX0/1/3 persist)."""
import os

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(__file__), "golden", "f1_text_mesh.npz")


def _host():
    import hydro_amd.mesh as mesh

    with np.load(FIX, allow_pickle=False) as z:
        raw = {k: (z[k].item() if z[k].ndim == 0 else z[k]) for k in z.files}
    return mesh._finish("F1_text", raw, np.float64)


def test_reference_f1_mesh_all_modes_vs_oracle(oracle):
    from hydro_amd.device import DeviceMesh

    host = _host()
    assert host.CELL == 6675 and 13 in np.unique(host.KLAS)
    orc = oracle.OracleRun(host)
    orc.step(200)
    res = {}
    for mode in ("two_kernel", "fused", "persistent"):
        dm = DeviceMesh(host, device=0)
        dm.step(200, mode)
        res[mode] = dm.fetch("H", "U", "V", "Z", "W")
        dm.close()
    for k in ("H", "U", "V", "Z", "W"):
        err = rel_err(res["two_kernel"][k], orc[k])
        print(f"{k}: rel {err:.2e}")
        assert err <= 1e-12, k
        for mode in ("fused", "persistent"):
            assert np.array_equal(res[mode][k], res["two_kernel"][k]), f"{mode} {k}"
    assert np.abs(res["two_kernel"]["U"]).max() > 0
