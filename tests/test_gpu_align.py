"""GPU: a dump written by hydro_dump after 100 fp64 steps of the channel
workload, two-kernel and fused (fluxes materialised) modes, passes the
alignment checker's fp64 gates against the CPU oracle's state."""
import os
import tempfile

import numpy as np
import pytest

from conftest import cached_mesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["two_kernel", "fused"])
def test_dump_passes_alignment_gates(oracle, mode):
    import hydro_amd
    from hydro_amd.align import compare, read_dump
    from hydro_amd.device import DeviceMesh

    host = cached_mesh("channel_6675")
    orc = oracle.OracleRun(host)
    orc.step(100)
    ref = {k: orc[k] for k in ("H", "U", "V", "Z", "W", "FLUX0", "FLUX1", "FLUX2", "FLUX3")}
    dm = DeviceMesh(host, device=0, tune=hydro_amd.default_tune(materialize_flux=1))
    dm.step(100, mode)
    path = os.path.join(tempfile.mkdtemp(), "gpu.bin")
    dm.dump(path)
    dm.close()
    got = read_dump(path)
    rep = compare(ref, got, host.AREA, host.KLAS, host.SIDE, "fp64", 100, host.NAC)
    print(mode, rep["verdict"], rep["reason"], {f: rep["fields"][f]["max_abs"] for f in ("H", "U", "V", "Z")})
    assert rep["verdict"] == "PASS", rep["reason"]


def test_step50_drift_inside_reference_envelope(oracle):
    """The 207,234-cell floodplain after 50 steps, device against the CPU
    oracle in the same precision (the reference's envelope compares two
    implementations at one precision): fp64 inside the fp64 gates, fp32
    inside the fp32 long-run gates (state p99 < 1e-3, conservation sums
    within 1e-5 relative)."""
    import hydro_amd
    from hydro_amd.align import compare
    from hydro_amd.device import DeviceMesh

    fields = ("H", "U", "V", "Z", "W", "FLUX0", "FLUX1", "FLUX2", "FLUX3")
    for dtype, prec in ((np.float64, "fp64"), (np.float32, "fp32")):
        host = cached_mesh("floodplain_207234", dtype=dtype)
        orc = oracle.OracleRun(host)
        orc.step(50)
        dm = DeviceMesh(host, device=0, tune=hydro_amd.default_tune(materialize_flux=1))
        dm.step(50, "fused")
        got = {k: v.astype(np.float64) for k, v in dm.fetch(*fields).items()}
        dm.close()
        ref = {k: orc[k].astype(np.float64) for k in fields}
        rep = compare(ref, got, host.AREA.astype(np.float64), host.KLAS, host.SIDE.astype(np.float64), prec, 50,
                      host.NAC)
        print(prec, rep["verdict"], rep["reason"], {f: rep["fields"][f]["max_abs"] for f in ("H", "U", "V", "Z")})
        assert rep["verdict"] == "PASS", rep["reason"]
