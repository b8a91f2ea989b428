"""The OUTPUT writer against vectors of the reference's own writer.

tests/golden/output_frames.npz holds, for a 230-cell mesh and three seeded
states, the five files that the reference's Python restatement of
MeshData::outputToFile (taichi_impl/output_writer.py; C++ original
cuda_native_impl/hydro-cal-src/src/mesh.cpp:535-839) writes, for DT = 4 and
DT = 0.5; tests/golden/make_output_golden.py made them.  The remaining
tests pin pieces against what the C++ fixes directly.
"""
import math
import os
import types

import numpy as np
import pytest

from hydro_amd import output

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "output_frames.npz")


def _golden():
    g = np.load(GOLDEN)
    arrays = {k: g[k] for k in ("COSF", "SINF", "SIDE", "AREA", "ZBC", "NAP", "XP", "YP", "XYMIN")}
    host = types.SimpleNamespace(arrays=arrays, CELL=int(g["CELL"]), HM1=float(g["HM1"]), DT=1.0, NDAYS=1)
    frames = [tuple(int(x) for x in f) for f in g["frames"]]
    states = [{k: g[f"state{i}_{k}"] for k in "HUVZ"} for i in range(len(frames))]
    return g, host, frames, states


@pytest.mark.parametrize("case", ["dt4", "dt05"])
def test_files_equal_reference_writer(tmp_path, case):
    g, host, frames, states = _golden()
    with output.OutputFiles(str(tmp_path), host, dt=float(g[f"{case}_DT"]), ndays=int(g[f"{case}_NDAYS"])) as w:
        for (jt, kt), s in zip(frames, states):
            w.frame(jt, kt, s)
    for fn in output.FILES:
        want = g[f"{case}_{fn}"].tobytes()
        got = (tmp_path / fn).read_bytes()
        if got != want:
            n = next((i for i in range(min(len(got), len(want))) if got[i] != want[i]), min(len(got), len(want)))
            pytest.fail(f"{fn}: first difference at byte {n}: got {got[max(0, n - 60):n + 40]!r} "
                        f"want {want[max(0, n - 60):n + 40]!r}")


def test_header_format_is_sticky(tmp_path):
    _, host, _, states = _golden()
    with output.OutputFiles(str(tmp_path), host, dt=4.0, ndays=2) as w:
        w.frame(0, 1, states[0])
        w.frame(1, 15, states[1])
    lines = [l for l in (tmp_path / "H2U2V2.OUT").read_text().split("\n") if l.startswith(" JT=")]
    assert lines == [" JT=    0  KT=    0  DT=  4 SEC  T= 0H  NSF= 0/0  WEC=0/  CQL=0  INE=0",
                     " JT=    2  KT=   14  DT=4.0000 SEC  T= 0H  NSF= 0/0  WEC=0/  CQL=0  INE=0"]
    assert (tmp_path / "TIMELOG.OUT").read_text() == "0.0000\n1.0000\n"
    # SIDE.OUT only at (0, 1)
    assert (tmp_path / "SIDE.OUT").read_text().count("COSF") == 1


def test_flow_angle_cases():
    """MeshData::FI: quadrants with pi = 3.1416, axes by the sequential
    tests (FI(0, 0) = 0, FI(0, -1) = 3 pi / 2), the 1e-9 snap."""
    pi = 3.1416
    x = np.array([1.0, -1.0, -1.0, 1.0, 0.0, 0.0, 0.0, 2.0, -2.0, 5e-10, 0.5])
    y = np.array([1.0, 1.0, -1.0, -1.0, 0.0, 3.0, -3.0, 0.0, 0.0, -1.0, -9e-10])
    want = np.array([math.pi / 4, pi - math.pi / 4, pi + math.pi / 4, 2 * pi - math.pi / 4, 0.0, pi / 2,
                     3 * pi / 2, 0.0, pi, 3 * pi / 2, 0.0]) * 57.298
    np.testing.assert_allclose(output.flow_angle(x, y), want, rtol=0, atol=1e-12)
    # the quadrant form against atan2 of the signed pair: the same angle up to the 3.1416 approximation
    rng = np.random.default_rng(5)
    u, v = rng.normal(size=1000), rng.normal(size=1000)
    ang = np.mod(np.arctan2(v, u), 2 * math.pi) * 57.298
    np.testing.assert_allclose(output.flow_angle(u, v), ang, rtol=0, atol=2 * 57.298 * abs(math.pi - pi) + 1e-9)


def test_block_cadence():
    """setw(10) fixed 4 values, a newline + 5 spaces before every tenth,
    another before every hundredth."""
    s = output._block("H2=", np.arange(205) * 1.0, True)
    lines = s.split("\n")
    assert lines[0] == "     " and lines[1] == "     H2="
    assert lines[2] == "     " + "".join("%10.4f" % v for v in range(10))
    assert lines[12] == "     " and lines[13].startswith("     " + "%10.4f" % 100)
    assert len([l for l in lines if l.strip() == ""]) == 3     # the leading line and the two hundreds
    assert output._block("U2=", [-2e-5, -1e-12], False) == "\n     \n     U2=\n        -0.0000   -0.0000"


def test_dry_cells_and_snap(tmp_path):
    """Dry cells print Z2 = ZBC and zeros; |v| < 1e-9 prints +0."""
    C = 3
    arrays = {"COSF": np.ones(12), "SINF": np.zeros(12), "SIDE": np.ones(12), "AREA": np.ones(3),
              "ZBC": np.array([5.0, 6.0, 7.0]), "NAP": np.arange(1, 13, dtype=np.int32),
              "XP": np.arange(12.0), "YP": np.arange(12.0), "XYMIN": np.array([10.0, 20.0])}
    host = types.SimpleNamespace(arrays=arrays, CELL=C, HM1=0.01, DT=1.0, NDAYS=1)
    st = {"H": np.array([0.01, 1.0, 2.0]), "U": np.array([3.0, -1e-10, 0.3]), "V": np.array([4.0, 0.0, -0.4]),
          "Z": np.array([9.0, -2e-10, 9.0])}
    with output.OutputFiles(str(tmp_path), host) as w:
        w.frame(0, 1, st)
    zuv = (tmp_path / "ZUV.OUT").read_text()
    row = lambda *v: "\n     " + "".join("%10.4f" % x for x in v)
    assert "Z2=" + row(5, 0, 9) in zuv
    assert "W2=" + row(0, 0, 0.5) in zuv
    h = (tmp_path / "H2U2V2.OUT").read_text()
    assert "U2=" + row(0, 0, 0.3) in h
    tec = (tmp_path / "XY-TEC.DAT").read_text().split("\n")
    assert tec[3].startswith("10.0000 11.0000 ")          # XP + XIMIN
    assert tec[5].startswith("20.0000 21.0000 ")
    assert tec[-4:] == ["1 2 3 4 ", "5 6 7 8 ", "9 10 11 12 ", ""]


class _OracleDevice:
    """Stands in for a DeviceMesh in run_with_output: steps the CPU oracle."""

    def __init__(self, oracle, host):
        self.host = host
        self.run = oracle.OracleRun(host)
        self.calls = []

    def step(self, n, mode):
        self.calls.append(n)
        self.run.step(n)

    def fetch(self, *names):
        return {k: self.run[k].copy() for k in names}


@pytest.mark.parametrize("steps,ntoutput,frames,calls", [
    (130, 1, [(0, 1), (0, 60), (1, 60), (2, 13)], [59, 59, 12]),
    (130, 2, [(0, 1), (1, 60), (2, 13)], [118, 12]),
    (59, 1, [(0, 1), (0, 60)], [59]),
    (0, 1, [(0, 1)], []),
])
def test_run_with_output_cadence(tmp_path, oracle, steps, ntoutput, frames, calls):
    """benchmark.cu:104-124: a frame at (0, 1), then after the last step of
    every ntoutput-th day and after the last step; the steps between frames
    run as one batch."""
    from hydro_amd.mesh import make_channel_mesh

    host = make_channel_mesh(cells=600, steps=130, DT=1.0, MDT=60)
    dev = _OracleDevice(oracle, host)
    assert output.run_with_output(dev, str(tmp_path / "a"), steps, ntoutput=ntoutput) == frames
    assert dev.calls == calls
    # the same frames written by stepping one step at a time
    run = oracle.OracleRun(host)
    with output.OutputFiles(str(tmp_path / "b"), host) as w:
        w.frame(0, 1, {k: run[k] for k in "HUVZ"})
        s = 0
        for jt, kt in frames[1:]:
            while s < jt * (host.steps_per_day - 1) + kt - 1:
                run.step(1)
                s += 1
            w.frame(jt, kt, {k: run[k] for k in "HUVZ"})
    for fn in output.FILES:
        assert (tmp_path / "a" / fn).read_bytes() == (tmp_path / "b" / fn).read_bytes(), fn


def test_compare_dirs(tmp_path):
    g, host, frames, states = _golden()
    c = int(np.nonzero((states[2]["H"] > host.HM1) & (np.abs(states[2]["U"]) > 1e-3))[0][0])
    for d, bump in (("a", 0.0), ("b", 0.25)):
        with output.OutputFiles(str(tmp_path / d), host, dt=4.0, ndays=3) as w:
            for i, ((jt, kt), s) in enumerate(zip(frames, states)):
                s = dict(s)
                if i == 2:
                    s["U"] = s["U"].copy()
                    s["U"][c] += bump
                w.frame(jt, kt, s)
    same = output.compare_dirs(str(tmp_path / "a"), str(tmp_path / "a"))
    assert same["all_byte_equal"] and len(same["blocks"]) == 3 * 6
    rep = output.compare_dirs(str(tmp_path / "a"), str(tmp_path / "b"))
    assert not rep["all_byte_equal"]
    assert rep["files"]["SIDE.OUT"]["byte_equal"] and not rep["files"]["H2U2V2.OUT"]["byte_equal"]
    u = rep["blocks"]["2/U2"]
    assert u["n_differ"] == 1 and abs(u["max_abs"] - 0.25) < 1.1e-4
    assert rep["blocks"]["1/U2"]["n_differ"] == 0 and rep["blocks"]["2/H2"]["n_differ"] == 0
    assert output.main([str(tmp_path / "a"), str(tmp_path / "a")]) == 0
