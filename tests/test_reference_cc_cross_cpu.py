"""Cross-check of the oracle against the reference CPU solver
(hydro-cal-fortan/calculate.cpp, built unmodified with a state-writing driver
into oracle/_ref/hydro_cc_state when the reference checkout is present), on
the branches where the two discretizations agree (no GPU).

The solver differs from variant A by design (two-sided Osher, per-cell CFL
step, +-5 m/s cap, calculate.cpp:113-256; it is the timed CPU baseline, not
the accuracy oracle), and it has two bugs of its own that this check steers
clear of: a neighbour whose 0-based id is 0 is taken for "no neighbour"
(`NC == 0`, calculate.cpp:124), which drains the cells next to cell 0, and
series paths that index past their arrays (so no discharge or level edges).

* Lake at rest in a walls-only basin, every cell wet: after one step depth
  and level equal the oracle's bit for bit (variants A and B) away from
  cell 0; velocities stay at rounding in both; after ten steps the depths
  agree to rounding.
* The same with dry cells (beds above the level, none shallow): every dry
  cell keeps H = HM1, Z = ZB1 in both, bit for bit, and the wet cells agree
  to rounding after one step.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import cached_mesh

EXE = os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "hydro_cc_state")
pytestmark = pytest.mark.skipif(not os.path.exists(EXE),
                                reason="oracle/_ref/hydro_cc_state not built (needs the reference checkout)")


def _basin(dry):
    host = cached_mesh("channel_6675").copy()
    a = host.arrays
    k = a["KLAS"]
    k[k != 0] = 4
    zbc = a["ZBC"].copy()
    if dry:
        level = float(np.median(zbc) + 0.6)
        shallow = (zbc > level - 0.05) & (zbc < level + 0.01)
        zbc[shallow] = level + 0.01            # dry, not shallow: no branch where the two differ
    else:
        level = float(zbc.max() + 0.5)
    a["ZBC"][:] = zbc
    a["ZB1"][:] = zbc + host.HM1
    wet = zbc < level
    a["H"][:] = np.where(wet, level - zbc, host.HM1)
    a["Z"][:] = np.where(wet, level, zbc + host.HM1)
    a["U"][:] = 0
    a["V"][:] = 0
    return host, level, wet


def _far_from_cell0(host, steps):
    """Cells more than steps + 1 edges away from cell 0 (the solver's
    NC == 0 bug spreads one ring per step)."""
    C = host.CELL
    nac = host.arrays["NAC"].reshape(C, 4)
    dist = np.full(C, np.iinfo(np.int64).max)
    dist[0] = 0
    front, d = [0], 0
    while front:
        d += 1
        nxt = []
        for c in front:
            for n in nac[c]:
                if n > 0 and dist[n - 1] > d:
                    dist[n - 1] = d
                    nxt.append(n - 1)
        front = nxt
    return dist > steps + 1


def _run_cc(host, level, wet, steps):
    from oracle import cc_input

    cc = host.copy()
    cc.arrays["Z"][:] = np.where(wet, level, host.ZBC)    # its initial level (dry where Z <= ZBC)
    root = tempfile.mkdtemp(prefix="hydro_cc_x_")
    cc_input.write_case(cc, root, steps)
    out = os.path.join(root, "state.bin")
    subprocess.run([EXE, out], cwd=root, check=True, stdout=subprocess.DEVNULL, timeout=300)
    raw = open(out, "rb").read()
    n, st = np.frombuffer(raw[:8], np.int32)
    assert n == host.CELL and st == steps
    H, U, V, Z = np.frombuffer(raw[8:], np.float64).reshape(4, n)
    return {"H": H, "U": U, "V": V, "Z": Z}


@pytest.mark.parametrize("variant", [0, 1], ids=["A", "B"])
def test_lake_at_rest_matches_reference_solver(oracle, variant):
    host, level, wet = _basin(dry=False)
    assert wet.all()
    for steps in (1, 10):
        cc = _run_cc(host, level, wet, steps)
        r = oracle.OracleRun(host, variant=variant)
        r.step(steps)
        far = _far_from_cell0(host, steps)
        assert far.sum() > 0.9 * host.CELL
        if steps == 1:
            assert np.array_equal(cc["H"][far], r["H"][far])
            assert np.array_equal(cc["Z"][far], r["Z"][far])
            assert np.array_equal(r["H"], host.H)                  # the oracle: invariant everywhere
        else:
            assert np.max(np.abs(cc["H"][far] - r["H"][far])) < 1e-14
        for k in ("U", "V"):
            assert np.max(np.abs(cc[k][far])) < 1e-13 and np.max(np.abs(r[k])) < 1e-13


def test_lake_with_dry_cells_matches_reference_solver(oracle):
    host, level, wet = _basin(dry=True)
    assert 0.1 * host.CELL < (~wet).sum() < 0.5 * host.CELL
    cc = _run_cc(host, level, wet, 1)
    r = oracle.OracleRun(host)
    r.step(1)
    far = _far_from_cell0(host, 1)
    dry = ~wet & far
    for k in ("H", "Z", "U", "V"):
        assert np.array_equal(cc[k][dry], r[k][dry]), k
    assert np.all(r["H"][~wet] == host.HM1)
    w = wet & far
    assert np.max(np.abs(cc["H"][w] - r["H"][w]) / r["H"][w]) < 1e-15
    assert np.array_equal(cc["Z"][w], r["Z"][w]) or np.max(np.abs(cc["Z"][w] - r["Z"][w])) < 1e-14
