"""GPU: OUTPUT files of a run on the device (run_with_output: a frame at
(0, 1), one at the end of every day, one after the last step) against the
same writer fed with the CPU oracle's states at those steps.  Same files
line for line; numbers within one unit of the last printed digit (the
device and the oracle agree to ~1e-12 relative, so a value on a rounding
boundary of the 4-decimal format may print one unit apart)."""
import numpy as np
import pytest

from hydro_amd import output

pytestmark = pytest.mark.gpu


def _compare(a, b):
    la, lb = a.split("\n"), b.split("\n")
    assert len(la) == len(lb)
    flips = 0
    for x, y in zip(la, lb):
        if x == y:
            continue
        tx, ty = x.split(), y.split()
        assert len(tx) == len(ty), (x, y)
        for s, t in zip(tx, ty):
            if s != t:
                assert abs(float(s) - float(t)) <= 1.0001e-4, (s, t)
                flips += 1
    return flips


@pytest.mark.parametrize("mode", ["fused", "two_kernel"])
def test_run_with_output_matches_oracle_frames(tmp_path, oracle, mode):
    from hydro_amd.device import DeviceMesh
    from hydro_amd.mesh import make_channel_mesh

    host = make_channel_mesh(cells=6675, steps=130, DT=1.0, MDT=60, rotate=0.3)
    steps = 130
    dm = DeviceMesh(host, device=0)
    frames = output.run_with_output(dm, str(tmp_path / "gpu"), steps, ntoutput=1, mode=mode)
    dm.close()
    assert frames == [(0, 1), (0, 60), (1, 60), (2, 13)]

    orc = oracle.OracleRun(host)
    done = 0
    with output.OutputFiles(str(tmp_path / "cpu"), host) as w:
        for jt, kt in frames:
            s = 0 if (jt, kt) == (0, 1) else jt * (host.steps_per_day - 1) + kt - 1
            if s > done:
                orc.step(s - done)
                done = s
            w.frame(jt, kt, {k: orc[k] for k in "HUVZ"})
    assert done == steps
    total = 0
    for fn in output.FILES:
        a = (tmp_path / "gpu" / fn).read_text()
        b = (tmp_path / "cpu" / fn).read_text()
        total += _compare(a, b)
    print(f"values one digit apart: {total}")
    assert total <= 10
