"""Reader of the reference's text mesh directories (the ``.DAT`` format).

Reads a directory laid out like the reference's ``data/*_text`` meshes (all
files in one directory) or like ``MeshData::InitFromFile`` expects
(``BINFOR/`` and ``SOURCES/`` subdirectories), and returns a
:class:`~hydro_amd.mesh.HostMesh` built the way the native loader builds its
arrays (cuda_native_impl/hydro-cal-src/src/mesh.cpp:5-521, expanded to cells
as cell.cpp:92-147 does):

* scalars, one per line with trailing comments: TIME (MDT, NDAYS), GIRD
  (NOD, CEL), DEPTH (HM1, HM2), BOUNDARY (NZ, NQ, NZQ, NHQ, ...), CALTIME
  (start stamp, DT), JL; a directory without TIME/CALTIME gets MDT = 3600,
  NDAYS = 50, DT = 1 (taichi_impl/mesh_loader_f1.py:80-93);
* per cell after a header line: PNAC, PNAP, PKLAS (``NO v0 v1 v2 v3``), PZBC,
  INITIALLEVEL, INITIALU1, INITIALV1, CV (Manning n); PXY per node
  (``NO X Y``);
* boundary entries: MBZ / MBQ / MBZQ (a count of series files, then
  ``NO cell group`` per boundary cell).

The series files under BOUNDE/ (NZ, NQ and NZQ) are read and turned into
the per-cell QT/DQT/ZT/DZT series and the NHQ1/ZW/QW rating curves by the
library's ``hydro_take_boundary`` (MeshData::take_boundary_for_two_d and
BOUNDRYinterp, mesh.cpp:374-533), after the library's
``hydro_precalculate`` (MeshData::preCalculate) has given the geometry, the
wet/dry initial state and the side lengths the rating curves use.  Weirs
(NWE > 0 with KLAS 6 slots) are not read; a mesh that needs them raises.
"""
from __future__ import annotations

import os

import numpy as np

from .mesh import HostMesh, _finish


def _path(root, name):
    for sub in ("", "BINFOR", "SOURCES"):
        p = os.path.join(root, sub, name)
        if os.path.exists(p):
            return p
    return None


def _first_tokens(path):
    out = []
    with open(path, "r", encoding="latin-1") as f:
        for line in f:
            parts = line.split()
            if parts:
                out.append(parts[0])
    return out


def _scalars(root, name, n, default=None):
    p = _path(root, name + ".DAT")
    if p is None:
        if default is None:
            raise FileNotFoundError(os.path.join(root, name + ".DAT"))
        return default
    toks = _first_tokens(p)
    return toks[:n]


def _table(root, name, rows, cols):
    """A header line, then `rows` lines whose first `cols` tokens are kept."""
    p = _path(root, name + ".DAT")
    if p is None:
        raise FileNotFoundError(os.path.join(root, name + ".DAT"))
    with open(p, "r", encoding="latin-1") as f:
        f.readline()
        data = []
        for line in f:
            parts = line.split()
            if not parts:
                continue
            data.append(parts[:cols])
            if len(data) == rows:
                break
    if len(data) != rows:
        raise ValueError(f"{p}: {len(data)} rows, expected {rows}")
    return np.array(data, dtype=np.float64)


def _boundary_cells(root, name, count):
    """MBZ / MBQ: (number of series files, [(cell 0-based, group)])."""
    p = _path(root, name + ".DAT")
    if p is None or count == 0:
        return 0, []
    with open(p, "r", encoding="latin-1") as f:
        nfiles = int(f.readline().split()[0])
        out = []
        for line in f:
            parts = line.split()
            if len(parts) >= 3:
                out.append((int(parts[1]) - 1, int(parts[2])))
            if len(out) == count:
                break
    return nfiles, out


def load_dat(root: str, dtype=np.float64, ndays=None, name=None) -> HostMesh:
    """Read a reference mesh directory.  `ndays` keeps only the first rows of
    the boundary series (the increments of the last kept row still come from
    the row after it, as in the full series)."""
    MDT, NDAYS = (int(float(x)) for x in _scalars(root, "TIME", 2, default=["3600", "50"]))
    NOD, C = (int(float(x)) for x in _scalars(root, "GIRD", 2))
    HM1, HM2 = (float(x) for x in _scalars(root, "DEPTH", 2))
    b = [int(float(x)) for x in _scalars(root, "BOUNDARY", 6)]
    NZ, NQ, NZQ = b[0], b[1], b[2]
    NWE = b[4] if len(b) > 4 else 0
    DT = float(_scalars(root, "CALTIME", 2, default=["0", "1"])[1])
    JL = float(_scalars(root, "JL", 1, default=["0"])[0])
    nac = _table(root, "PNAC", C, 5).astype(np.int64)[:, 1:]
    nap = _table(root, "PNAP", C, 5).astype(np.int64)[:, 1:]
    klas = _table(root, "PKLAS", C, 5).astype(np.int64)[:, 1:]
    if np.any(klas == 6) and NWE > 0:
        raise NotImplementedError("weir boundaries (KLAS 6) are not read")
    zbc = _table(root, "PZBC", C, 1)[:, 0]
    xy = _table(root, "PXY", NOD, 3)
    z0 = _table(root, "INITIALLEVEL", C, 1)[:, 0]
    u0 = _table(root, "INITIALU1", C, 1)[:, 0]
    v0 = _table(root, "INITIALV1", C, 1)[:, 0]
    n = _table(root, "CV", C, 1)[:, 0]

    nd = NDAYS if ndays is None else max(1, min(int(ndays), NDAYS))
    NHQ = b[3] if len(b) > 3 else 0

    def entries(name, count):
        nfiles, ent = _boundary_cells(root, name, count)
        cell = np.array([c + 1 for c, _ in ent], dtype=np.int32)
        group = np.array([g for _, g in ent], dtype=np.int32)
        return nfiles, cell, group

    # the series and rating curves themselves are read by the library
    # (hydro_take_boundary) once preCalculate has given the side lengths
    bounde = dict(dir=os.path.join(root, "BOUNDE"), NDAYS=NDAYS, NHQ=NHQ,
                  NZ=entries("MBZ", NZ), NQ=entries("MBQ", NQ), NZQ=entries("MBZQ", NZQ))
    raw = dict(CELL=C, NOD=NOD, XP=xy[:, 1].copy(), YP=xy[:, 2].copy(), ZBC=zbc, Zinit=z0, n=n, U0=u0, V0=v0,
               NAP=nap.reshape(-1).astype(np.int32), NAC=nac.reshape(-1).astype(np.int32),
               KLAS=klas.reshape(-1).astype(np.int32), NDAYS=nd, MDT=MDT, DT=DT, HM1=HM1, HM2=HM2, JL=JL,
               NHQ=NHQ, bounde=bounde)
    nv = np.where((nap[:, 3] == 0) | (nap[:, 3] == nap[:, 0]), 3, 4)
    idx = np.where(nap > 0, nap - 1, 0)
    w = (np.arange(4)[None, :] < nv[:, None]).astype(np.float64)
    raw["xc"] = (xy[idx, 1] * w).sum(1) / nv
    raw["yc"] = (xy[idx, 2] * w).sum(1) / nv
    return _finish(name or os.path.basename(os.path.normpath(root)), raw, dtype)
