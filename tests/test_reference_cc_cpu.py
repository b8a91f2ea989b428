"""The reference CPU solver (hydro-cal-fortan/calculate.cpp, built unmodified
into oracle/_ref/hydro_cc when the reference checkout is present) runs on the
input directory oracle/cc_input.py writes for a synthetic mesh: the CPU
baseline bench.py reports."""
import os
import subprocess
import tempfile

import pytest

from conftest import cached_mesh

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                                                "hydro_cc")),
                                reason="oracle/_ref/hydro_cc not built (needs the reference checkout)")


def test_reference_solver_runs_on_written_case():
    from oracle import cc_input

    host = cached_mesh("channel_6675")
    root = tempfile.mkdtemp(prefix="hydro_cc_test_")
    d = cc_input.write_case(host, root, steps=3)
    for name in ("TIME", "GIRD", "DEPTH", "BOUNDARY", "CALTIME", "INITIALLEVEL", "INITIALU1", "INITIALV1", "CV"):
        assert os.path.exists(os.path.join(d, name + ".DAT")), name
    with open(os.path.join(d, "GIRD.DAT")) as f:
        assert [int(x) for x in f.read().split()] == [len(host.arrays["XP"]), host.CELL]
    r = subprocess.run([cc_input.CC_PATH], cwd=root, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    # one line per simulated day (NDAYS = 1): "0"
    assert r.stdout.split()[-1] == "0"
