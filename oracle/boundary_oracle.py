"""Boundary series and rating curves of a .DAT directory, restated in Python:
MeshData::take_boundary_for_two_d and BOUNDRYinterp
(cuda_native_impl/hydro-cal-src/src/mesh.cpp:374-533) and the expansion of
boundary entries to cells in CellView's constructor (cell.cpp:92-126).

TEST INFRASTRUCTURE ONLY: the checker of the library's hydro_take_boundary,
imported by tests/.  Pure-Python loops over small inputs.

Arithmetic follows the reference's types: Real is numpy.float64 (double) or
numpy.float32 (float) scalars, which divide by zero as IEEE does; the float
temporaries of mesh.cpp:405 and :441 are numpy.float32, and `x / n` with an
int n stays in the type of x.  Numbers
are parsed as double and then rounded to Real; for an fp32 check, write
numbers that are exact in float, so that this equals the reference's direct
parse to float.
"""
from __future__ import annotations

import os

import numpy as np


def _real(dtype):
    return np.float64 if np.dtype(dtype) == np.float64 else np.float32


def read_series_file(path, real):
    """mesh.cpp:389-400 (and :421-437, :478-492): the count on the first
    line, then per line two `>>` reads; what a line lacks reads as 0."""
    with open(path, "r", encoding="latin-1") as f:
        lines = f.read().split("\n")
    n = int(lines[0].split()[0])
    first, second = [], []
    for i in range(n):
        toks = lines[1 + i].split() if 1 + i < len(lines) else []
        first.append(real(float(toks[0])) if len(toks) > 0 else real(0))
        second.append(real(float(toks[1])) if len(toks) > 1 else real(0))
    return first, second


def boundary_interp(t, times, vals, real):
    """BOUNDRYinterp (mesh.cpp:523-533): of the brackets [times[i-1],
    times[i]] that hold t, the last one gives the line through its ends; 0
    when none does.  (The reference's loop also tests a bracket ending one
    entry past the table, which reads outside it; not restated.)"""
    result = real(0)
    for i in range(1, len(times)):
        if t >= times[i - 1] and t <= times[i]:
            result = vals[i - 1] + (vals[i] - vals[i - 1]) / (times[i] - times[i - 1]) * (t - times[i - 1])
    return result


def _entry_rows(bounde_dir, kind, nfiles, groups, nrow, MDT, real, split):
    """Rows 0..nrow-1 of every entry (mesh.cpp:383-447)."""
    rows = [[real(0)] * nrow for _ in groups]
    for k in range(1, nfiles + 1):
        times, vals = read_series_file(os.path.join(bounde_dir, kind, f"{kind}{k:04d}.DAT"), real)
        kl = sum(1 for g in groups if g == k)                       # mesh.cpp:427-430
        for i in range(nrow):
            stime1 = real(0.0 + i / (24.0 * 3600.0 / MDT))          # mesh.cpp:394
            tmp = np.float32(boundary_interp(stime1, times, vals, real))   # float ZTTEMP / QTTEMP
            v = real(tmp / kl) if split else real(tmp)              # mesh.cpp:445: QTTEMP / kl in float
            for j, g in enumerate(groups):
                if g == k:
                    rows[j][i] = v
    return rows


def take_boundary(bounde_dir, CELL, NDAYS, rows, MDT, DT, HM2, NHQ, nz, nq, nzq, ZBC=None, SIDE=None, KLAS=None,
                  dtype=np.float64):
    """Per-cell series QT, DQT, ZT, DZT [rows*CELL] (column-major) and rating
    curves NHQ1 [CELL], ZW, QW [CELL*NHQ].  nz, nq, nzq: (number of files,
    1-based cells, file numbers) of the MBZ, MBQ and MBZQ entries."""
    real = _real(dtype)
    nrow = min(rows + 1, NDAYS)
    K0 = int(real(MDT) / real(DT))                                  # mesh.cpp:449
    out = {k: np.zeros(rows * CELL, dtype=dtype) for k in ("QT", "DQT", "ZT", "DZT")}
    for (nfiles, cells, groups), V, D, split, kind in ((nz, "ZT", "DZT", False, "NZ"), (nq, "QT", "DQT", True, "NQ")):
        per_entry = _entry_rows(bounde_dir, kind, nfiles, list(groups), nrow, MDT, real, split)
        for j, cell in enumerate(cells):                            # cell.cpp:110-126
            pos = int(cell) - 1
            for jt in range(rows):
                out[V][jt * CELL + pos] = per_entry[j][jt]
                if jt != NDAYS - 1:                                 # mesh.cpp:451-463
                    out[D][jt * CELL + pos] = (per_entry[j][jt + 1] - per_entry[j][jt]) / K0
                else:
                    out[D][jt * CELL + pos] = 0
    out["NHQ1"] = np.zeros(CELL, dtype=np.int32)
    out["ZW"] = np.zeros(max(CELL * NHQ, 1), dtype=dtype)
    out["QW"] = np.zeros(max(CELL * NHQ, 1), dtype=dtype)
    nfiles, cells, groups = nzq
    nent = len(cells)
    ZW = [[real(0)] * nent for _ in range(NHQ)]                     # cells.ZW[point][entry], mesh.cpp:470
    QW = [[real(0)] * nent for _ in range(NHQ)]
    NHQ1 = [0] * nent
    for k in range(1, nfiles + 1):                                  # mesh.cpp:472-519
        levels, flows = read_series_file(os.path.join(bounde_dir, "NZQ", f"NZQ{k:04d}.DAT"), real)
        if len(levels) > NHQ:
            raise ValueError("rating curve longer than NHQ")
        AQZH0 = [real(0)] * nent
        AQZH1 = [real(0)] * NHQ
        for i in range(len(levels)):
            for j in range(nent):
                if groups[j] != k:
                    continue
                c = int(cells[j]) - 1
                NHQ1[j] = len(levels)
                for jj in range(4):
                    if KLAS[4 * c + jj] == 3:
                        AQZH0[j] = real(np.fmax(levels[i] - real(ZBC[c]), real(HM2))) * real(SIDE[4 * c + jj])
                AQZH1[i] = AQZH1[i] + AQZH0[j]
            for j in range(nent):
                if groups[j] == k:
                    ZW[i][j] = levels[i]
                    QW[i][j] = flows[i] * AQZH0[j] / AQZH1[i]
    for j in range(nent):                                           # cell.cpp:103-109
        pos = int(cells[j]) - 1
        for i in range(NHQ):
            out["ZW"][pos * NHQ + i] = ZW[i][j]
            out["QW"][pos * NHQ + i] = QW[i][j]
        out["NHQ1"][pos] = NHQ1[j]
    return out
