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
