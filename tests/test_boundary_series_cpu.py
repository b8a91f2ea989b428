"""CPU: boundary series and rating curves of a .DAT directory.

The library's hydro_take_boundary (MeshData::take_boundary_for_two_d and
BOUNDRYinterp, cuda_native_impl/hydro-cal-src/src/mesh.cpp:374-533, with the
expansion to cells of cell.cpp:92-126) is compared bit for bit with the
oracle's restatement (oracle/boundary_oracle.py), in fp64 and fp32, on
synthetic BOUNDE directories written here.  Both are pinned by what the
definitions fix on their own:

* a row at a knot time is the knot's value (the last bracket holding a time
  wins), a row between knots lies on the line, a row outside the table is 0;
* every row goes through a float temporary, also in fp64;
* a discharge is split over the entries of its file in float;
* D * K0 recovers the row difference, and the last row of the series has
  D = 0, while a truncated series still takes its last increment from the
  row after it;
* the later of two entries of one cell wins;
* a rating curve with one cell hands it the whole discharge; on a curve of
  two cells the shares follow the wetted widths (depth above the bed, at
  least HM2, times the side of the last KLAS 3 slot) and sum to the whole;
* the oracle's interpolation equals BOUNDRYinterp of the reference's own
  compiled CPU solver (oracle/_ref/hydro_cc_interp, built from
  hydro-cal-fortan/calculate.cpp when the checkout is present).
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from hydro_amd import _lib
from oracle import boundary_oracle as bo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CC_INTERP = os.path.join(ROOT, "oracle", "_ref", "hydro_cc_interp")

CELL = 12
NHQ = 4
MDT = 21600            # 6 h per row: row i sits at i / 4 days
DT = 0.5               # K0 = 43200
HM2 = 0.25


def _write(path, pairs, count=None):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write(f"{len(pairs) if count is None else count}\n")
        for p in pairs:
            f.write("\t".join(str(x) for x in p) + "\n")


@pytest.fixture(scope="module")
def bounde(tmp_path_factory):
    """A BOUNDE directory: 3 level files, 2 discharge files, 2 rating
    curves.  Every number is exact in float."""
    d = str(tmp_path_factory.mktemp("bounde"))
    # levels: knots at 0, 0.5, 1.25 days, then a step up to 1.75
    _write(f"{d}/NZ/NZ0001.DAT", [(0, 2.5), (0.5, 3.0), (1.25, 1.5), (1.75, 4.0)])
    # starts after row 0: row 0 reads 0
    _write(f"{d}/NZ/NZ0002.DAT", [(0.125, 7.0), (3.0, 8.0)])
    # a value that is not exact in float (0.1): the float temporary rounds it
    _write(f"{d}/NZ/NZ0003.DAT", [(0, 0.1), (4.0, 0.1)])
    # discharges: file 1 split over 3 entries, file 2 over 1; a short line reads as 0
    _write(f"{d}/NQ/NQ0001.DAT", [(0, 1.0), (1.0, 5.0), (2.0, 0.5)])
    _write(f"{d}/NQ/NQ0002.DAT", [(0, 100.0), (0.75,), (1.0, 40.0)])
    # rating curves: levels, discharges
    _write(f"{d}/NZQ/NZQ0001.DAT", [(1.0, 10.0), (2.0, 30.0), (3.5, 70.0)])
    _write(f"{d}/NZQ/NZQ0002.DAT", [(0.5, 4.0), (1.5, 12.0)])
    return d


def _entries():
    nz = (3, [1, 2, 5, 2], [1, 1, 2, 3])   # cell 2: file 1, then file 3 (later wins)
    nq = (2, [7, 8, 9, 10], [1, 1, 1, 2])
    nzq = (2, [3, 4, 11], [1, 2, 2])
    return nz, nq, nzq


def _statics(dtype):
    ZBC = np.zeros(CELL, dtype=dtype)
    ZBC[2], ZBC[3], ZBC[10] = 0.5, 1.0, 0.0
    SIDE = np.full(4 * CELL, 2.0, dtype=dtype)
    SIDE[4 * 3 + 1] = 4.0
    SIDE[4 * 10 + 2] = 1.0
    KLAS = np.zeros(4 * CELL, dtype=np.int32)
    KLAS[4 * 2 + 0] = 3          # cell 3: slot 0
    KLAS[4 * 3 + 1] = 3          # cell 4: slot 1 (side 4)
    KLAS[4 * 10 + 0] = 3         # cell 11: slots 0 and 2, the last (side 1) counts
    KLAS[4 * 10 + 2] = 3
    return ZBC, SIDE, KLAS


def _native(lib, d, dtype, NDAYS, rows, nz, nq, nzq, statics):
    ZBC, SIDE, KLAS = statics
    keep = []

    def ia(a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        keep.append(a)
        return a.ctypes.data_as(ctypes.c_void_p) if a.size else None

    desc = _lib.BoundaryDesc(dtype=_lib.F64 if dtype == np.float64 else _lib.F32, CELL=CELL, NDAYS=NDAYS, rows=rows,
                             MDT=MDT, NHQ=NHQ, DT=DT, HM2=HM2)
    for key, (nf, cells, groups) in (("Z", nz), ("Q", nq), ("ZQ", nzq)):
        setattr(desc, "N" + key, len(cells))
        setattr(desc, "NN" + key + "0", nf)
        setattr(desc, "MB" + key, ia(cells))
        setattr(desc, "NN" + key, ia(groups))
    desc.ZBC = ZBC.ctypes.data_as(ctypes.c_void_p)
    desc.SIDE = SIDE.ctypes.data_as(ctypes.c_void_p)
    desc.KLAS = KLAS.ctypes.data_as(ctypes.c_void_p)
    out = {k: np.full(rows * CELL, 9.0, dtype=dtype) for k in ("QT", "DQT", "ZT", "DZT")}   # overwritten in full
    out["NHQ1"] = np.full(CELL, 9, dtype=np.int32)
    out["ZW"] = np.full(CELL * NHQ, 9.0, dtype=dtype)
    out["QW"] = np.full(CELL * NHQ, 9.0, dtype=dtype)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    st = lib.hydro_take_boundary(ctypes.byref(desc), d.encode(), p(out["QT"]), p(out["DQT"]), p(out["ZT"]),
                                 p(out["DZT"]), p(out["NHQ1"]), p(out["ZW"]), p(out["QW"]))
    return st, out


@pytest.fixture(scope="module")
def lib():
    return _lib.load()


def _oracle(d, dtype, NDAYS, rows, statics):
    nz, nq, nzq = _entries()
    ZBC, SIDE, KLAS = statics
    with np.errstate(divide="ignore", invalid="ignore"):
        return bo.take_boundary(d, CELL, NDAYS, rows, MDT, DT, HM2, NHQ, nz, nq, nzq, ZBC, SIDE, KLAS, dtype=dtype)


def _row(a, jt, cell):
    return a[jt * CELL + cell - 1]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("NDAYS,rows", [(10, 10), (10, 4)])
def test_native_equals_oracle(lib, bounde, dtype, NDAYS, rows):
    statics = _statics(dtype)
    st, got = _native(lib, bounde, dtype, NDAYS, rows, *_entries(), statics)
    assert st == _lib.OK, _lib.last_error(lib)
    ref = _oracle(bounde, dtype, NDAYS, rows, statics)
    for k in ("QT", "DQT", "ZT", "DZT", "NHQ1", "ZW", "QW"):
        assert got[k].dtype == ref[k].dtype, k
        assert np.array_equal(got[k].view(np.uint8), ref[k].view(np.uint8)), k


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("who", ["native", "oracle"])
def test_series_closed_forms(lib, bounde, dtype, who):
    NDAYS, rows = 10, 10
    statics = _statics(dtype)
    if who == "native":
        st, a = _native(lib, bounde, dtype, NDAYS, rows, *_entries(), statics)
        assert st == _lib.OK
    else:
        a = _oracle(bounde, dtype, NDAYS, rows, statics)
    ZT, DZT, QT, DQT = a["ZT"], a["DZT"], a["QT"], a["DQT"]
    K0 = int(MDT / DT)
    # cell 1, file 1: knots at rows 0 (2.5), 2 (3.0), 5 (1.5), 7 (4.0); the
    # row at a knot is the knot value, between knots the line, past the table 0
    assert [float(_row(ZT, r, 1)) for r in (0, 2, 5, 7)] == [2.5, 3.0, 1.5, 4.0]
    assert float(_row(ZT, 1, 1)) == 2.75
    assert float(_row(ZT, 6, 1)) == 2.75
    assert [float(_row(ZT, r, 1)) for r in (8, 9)] == [0.0, 0.0]
    # cell 2: file 1 first, then file 3 -- the later entry wins; file 3 holds
    # 0.1, which comes out rounded through float even in fp64
    assert _row(ZT, 0, 2) == dtype(np.float32(0.1))
    # cell 5, file 2 starts at 0.125 days: row 0 lies before it
    assert float(_row(ZT, 0, 5)) == 0.0 and float(_row(ZT, 1, 5)) > 7.0
    # increments: D * K0 is the row difference, D = 0 on row NDAYS - 1
    for c in (1, 2, 5):
        for jt in range(rows - 1):
            assert _row(DZT, jt, c) == (_row(ZT, jt + 1, c) - _row(ZT, jt, c)) / dtype(K0)
        assert _row(DZT, rows - 1, c) == 0
    # discharges: file 1 over 3 entries, in float; file 2 whole
    third = np.float32(np.float32(1.0) / np.float32(3))
    assert _row(QT, 0, 7) == dtype(third) and _row(QT, 0, 9) == dtype(third)
    assert float(_row(QT, 0, 10)) == 100.0
    # NQ0002's second line lacks its value: (0.75, 0); the row at 0.5 days
    # lies on the line from (0, 100) to (0.75, 0), evaluated in Real, then
    # rounded through float
    R = dtype
    assert _row(QT, 2, 10) == R(np.float32(R(100.0) + (R(0) - R(100.0)) / R(0.75) * R(0.5)))
    # cells with no entry stay 0
    for c in (3, 4, 6, 11, 12):
        assert all(float(_row(ZT, r, c)) == 0 and float(_row(QT, r, c)) == 0 for r in range(rows))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("who", ["native", "oracle"])
def test_rating_curve_closed_forms(lib, bounde, dtype, who):
    statics = _statics(dtype)
    if who == "native":
        st, a = _native(lib, bounde, dtype, 10, 10, *_entries(), statics)
        assert st == _lib.OK
    else:
        a = _oracle(bounde, dtype, 10, 10, statics)
    ZW = a["ZW"].reshape(CELL, NHQ)
    QW = a["QW"].reshape(CELL, NHQ)
    NHQ1 = a["NHQ1"]
    # curve 1 holds cell 3 alone: the whole discharge, at every point
    assert NHQ1[2] == 3
    assert list(ZW[2, :3]) == [1.0, 2.0, 3.5] and list(QW[2, :3]) == [10.0, 30.0, 70.0]
    assert ZW[2, 3] == 0 and QW[2, 3] == 0
    # curve 2: cells 4 (bed 1.0, side 4) and 11 (bed 0, last KLAS 3 side 1)
    assert NHQ1[3] == 2 and NHQ1[10] == 2
    for i, (z, q) in enumerate([(0.5, 4.0), (1.5, 12.0)]):
        w4 = max(z - 1.0, HM2) * 4.0
        w11 = max(z - 0.0, HM2) * 1.0
        assert ZW[3, i] == z and ZW[10, i] == z
        assert abs(float(QW[3, i]) - q * w4 / (w4 + w11)) <= 4 * np.finfo(dtype).eps * q
        assert abs(float(QW[10, i]) - q * w11 / (w4 + w11)) <= 4 * np.finfo(dtype).eps * q
        assert abs(float(QW[3, i]) + float(QW[10, i]) - q) <= 4 * np.finfo(dtype).eps * q
    # cells on no curve
    for c in (0, 1, 4, 5, 6, 7, 8, 9, 11):
        assert NHQ1[c] == 0 and not ZW[c].any() and not QW[c].any()


def test_truncated_series_keeps_last_increment(lib, bounde):
    full = _native(lib, bounde, np.float64, 10, 10, *_entries(), _statics(np.float64))[1]
    part = _native(lib, bounde, np.float64, 10, 4, *_entries(), _statics(np.float64))[1]
    for k in ("QT", "DQT", "ZT", "DZT"):
        assert np.array_equal(part[k], full[k][: 4 * CELL]), k
    assert np.any(part["DZT"][3 * CELL:] != 0)


def test_errors(lib, bounde, tmp_path):
    nz, nq, nzq = _entries()
    s = _statics(np.float64)
    st, _ = _native(lib, str(tmp_path), np.float64, 10, 10, nz, nq, nzq, s)          # no files
    assert st == _lib.ERR_INVALID and b"cannot open" in lib.hydro_last_error()
    st, _ = _native(lib, bounde, np.float64, 10, 10, (3, [1, 13], [1, 2]), nq, nzq, s)  # cell 13 of 12
    assert st == _lib.ERR_INVALID and b"outside" in lib.hydro_last_error()
    st, _ = _native(lib, bounde, np.float64, 10, 11, nz, nq, nzq, s)                  # rows > NDAYS
    assert st == _lib.ERR_INVALID
    _write(str(tmp_path / "NZQ" / "NZQ0001.DAT"), [(1, 1)] * (NHQ + 1))
    st, _ = _native(lib, str(tmp_path), np.float64, 10, 10, (0, [], []), (0, [], []), (1, [3], [1]), s)
    assert st == _lib.ERR_INVALID and b"more than NHQ" in lib.hydro_last_error()


def test_boundary_increments(lib):
    rng = np.random.default_rng(5)
    for dtype in (np.float64, np.float32):
        V = rng.standard_normal(6 * 7).astype(dtype)
        D = np.full_like(V, 3.0)
        st = lib.hydro_boundary_increments(_lib.F64 if dtype == np.float64 else _lib.F32, 7, 6, 6, 3,
                                           V.ctypes.data_as(ctypes.c_void_p), D.ctypes.data_as(ctypes.c_void_p))
        assert st == _lib.OK
        V2, D2 = V.reshape(6, 7), D.reshape(6, 7)
        assert np.array_equal(D2[:-1], (V2[1:] - V2[:-1]) / dtype(3))
        assert not D2[-1].any()


@pytest.mark.skipif(not os.path.exists(CC_INTERP), reason="reference CPU solver not built (no checkout)")
def test_oracle_interp_equals_reference_solver():
    rng = np.random.default_rng(11)
    times = np.cumsum(rng.uniform(0.05, 1.0, 9))
    times[0] = 0.0
    vals = rng.uniform(-3, 9, 9)
    # knots, their neighbours, random points and gaps; all below the last knot
    q = np.concatenate([times[:-1], np.nextafter(times[1:-1], -np.inf), rng.uniform(0, times[-1], 40),
                        [-0.5]])
    inp = f"{len(times)}\n" + "".join(f"{float(t)!r} {float(v)!r}\n" for t, v in zip(times, vals)) + \
        f"{len(q)}\n" + "".join(f"{float(x)!r}\n" for x in q)
    out = subprocess.run([CC_INTERP], input=inp, capture_output=True, text=True, check=True).stdout.split()
    ref = [float(x) for x in out]
    got = [float(bo.boundary_interp(np.float64(x), list(times), list(vals), np.float64)) for x in q]
    assert got == ref


def _mesh_dir(root):
    """Two unit-square cells side by side, a level series on cell 1 and a
    rating curve on cell 2's KLAS 3 slot, in the reference's .DAT layout."""
    def put(name, text):
        with open(os.path.join(root, name), "w") as f:
            f.write(text)

    put("TIME.DAT", "21600 !MDT\n4 !NDAYS\n")
    put("GIRD.DAT", "6 !NOD\n2 !CEL\n")
    put("DEPTH.DAT", "0.001 !HM1\n0.25 !HM2\n")
    put("BOUNDARY.DAT", "1 !NZ\n0 !NQ\n1 !NZQ\n3 !NHQ\n0 !NWE\n0 !NDI\n")
    put("CALTIME.DAT", "0 !start\n0.5 !DT\n")
    put("PXY.DAT", "NO X Y\n1 0 0\n2 1 0\n3 2 0\n4 0 1\n5 1 1\n6 2 1\n")
    put("PNAP.DAT", "NO 1 2 3 4\n1 1 2 5 4\n2 2 3 6 5\n")
    put("PNAC.DAT", "NO 1 2 3 4\n1 0 2 0 0\n2 0 0 0 1\n")
    put("PKLAS.DAT", "NO 1 2 3 4\n1 4 0 4 1\n2 4 3 4 0\n")
    put("PZBC.DAT", "NO ZBC\n1 0.0\n2 0.5\n")
    put("INITIALLEVEL.DAT", "NO Z\n1 1.0\n2 1.0\n")
    put("INITIALU1.DAT", "NO U\n1 0\n2 0\n")
    put("INITIALV1.DAT", "NO V\n1 0\n2 0\n")
    put("CV.DAT", "NO n\n1 0.03\n2 0.03\n")
    put("MBZ.DAT", "1\n1 1 1\n")
    put("MBQ.DAT", "0\n")
    put("MBZQ.DAT", "1\n1 2 1\n")
    _write(os.path.join(root, "BOUNDE", "NZ", "NZ0001.DAT"), [(0, 1.0), (1.0, 2.0)])
    _write(os.path.join(root, "BOUNDE", "NZQ", "NZQ0001.DAT"), [(0.5, 0.0), (1.5, 8.0), (2.5, 24.0)])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_load_dat_reads_series_and_rating_curves(tmp_path, dtype):
    from hydro_amd.dat import load_dat

    _mesh_dir(str(tmp_path))
    m = load_dat(str(tmp_path), dtype=dtype)
    a = m.arrays
    assert m.NHQ == 3 and m.NDAYS == 4 and m.steps_per_day == 43200
    # level series of cell 1: rows at 0, 0.25, 0.5, 0.75 days on the line 1 + t
    assert list(a["ZT"].reshape(4, 2)[:, 0]) == [1.0, 1.25, 1.5, 1.75]
    assert not a["ZT"].reshape(4, 2)[:, 1].any()
    # cell 2 alone on its curve: the whole discharge at every level
    assert a["NHQ1"][1] == 3 and a["NHQ1"][0] == 0
    assert list(a["ZW"].reshape(2, 3)[1]) == [0.5, 1.5, 2.5]
    assert list(a["QW"].reshape(2, 3)[1]) == [0.0, 8.0, 24.0]
    # the side the widths came from is preCalculate's: unit squares
    assert np.all(a["SIDE"] == 1)
