"""CPU: the alignment checker (hydro_amd.align) — verdicts on identical and
perturbed states, exact conservation sums, dump round trip."""
import math
import os
import tempfile
from fractions import Fraction

import numpy as np

from conftest import cached_mesh


def _state(oracle, steps=20, dtype=np.float64):
    host = cached_mesh("channel_6675", dtype=dtype)
    orc = oracle.OracleRun(host)
    orc.step(steps)
    s = {k: orc[k].copy() for k in ("H", "U", "V", "Z", "W", "FLUX0", "FLUX1", "FLUX2", "FLUX3")}
    s.update(SLCOS=host.SLCOS, SLSIN=host.SLSIN, SIDE=host.SIDE)
    return host, s


def test_identical_states_pass(oracle):
    from hydro_amd.align import compare

    host, s = _state(oracle)
    rep = compare(s, s, host.AREA, host.KLAS, host.SIDE, "fp64", 20, host.NAC)
    assert rep["verdict"] == "PASS", rep["reason"]
    assert all(rep["fields"][f]["bit_exact_frac"] == 1.0 for f in ("H", "U", "V", "Z", "W"))
    assert rep["conservation"]["klas10_inflow"]["ref"] != 0      # the channel's inflow edges count


def test_perturbations_fail_with_the_right_reason(oracle):
    from hydro_amd.align import compare

    host, s = _state(oracle)
    bad = dict(s)
    bad["H"] = s["H"].copy()
    bad["H"][123] += 1e-8
    rep = compare(s, bad, host.AREA, host.KLAS, host.SIDE, "fp64", 20, host.NAC)
    assert rep["verdict"] == "FAIL" and rep["reason"].startswith("H.max_abs")
    assert rep["fields"]["H"]["worst_cells"][0]["cell"] == 123
    # a 1e-10 shift on 5% of the cells: max_abs passes, p99 and mass do not
    bad["H"] = s["H"].copy()
    idx = np.arange(0, host.CELL, 20)
    bad["H"][idx] += 1e-10
    rep = compare(s, bad, host.AREA, host.KLAS, host.SIDE, "fp64", 20)
    assert rep["verdict"] == "FAIL"
    assert rep["conservation"]["mass"]["rel_diff"] > 1e-12
    bad["W"] = np.where(np.arange(host.CELL) == 7, np.nan, s["W"])
    bad["H"] = s["H"]
    rep = compare(s, bad, host.AREA, host.KLAS, host.SIDE, "fp64", 20)
    assert rep["verdict"] == "FAIL" and "NaN" in rep["reason"]


def test_fp32_step1_gates(oracle):
    from hydro_amd.align import compare

    host, s = _state(oracle, steps=1, dtype=np.float32)
    rep = compare(s, s, host.AREA, host.KLAS, host.SIDE, "fp32", 1)
    assert rep["verdict"] == "PASS", rep["reason"]
    bad = dict(s)
    bad["U"] = s["U"].copy()
    bad["U"][5] += np.float32(1e-5)
    rep = compare(s, bad, host.AREA, host.KLAS, host.SIDE, "fp32", 1)
    assert rep["verdict"] == "FAIL" and rep["reason"].startswith("U.max_abs")


def test_conservation_sums_are_exact():
    from hydro_amd.align import conservation

    rng = np.random.default_rng(3)
    H = rng.uniform(0, 5, 997)
    a = 10.0 ** rng.uniform(-3, 6, 997)
    st = {"H": H, "U": np.zeros_like(H), "V": np.zeros_like(H), "F0": np.zeros(4 * 997)}
    got = conservation(st, a, np.zeros(4 * 997, dtype=np.int32), np.ones(4 * 997))["mass"]
    exact = sum(Fraction(float(x)) for x in H * a)
    assert got == float(exact)
    assert got == math.fsum(H * a)


def test_dump_round_trip(oracle):
    from hydro_amd.align import read_dump, write_dump

    host, s = _state(oracle, steps=3)
    path = os.path.join(tempfile.mkdtemp(), "state.bin")
    write_dump(path, s)
    r = read_dump(path)
    for k in ("H", "U", "V", "Z", "SLCOS", "SLSIN", "SIDE"):
        assert np.array_equal(r[k], s[k]), k
    for i in range(4):
        assert np.array_equal(r[f"F{i}"], s[f"FLUX{i}"])
    assert os.path.getsize(path) == 4 + 8 * (4 * host.CELL + 7 * 4 * host.CELL)
