"""Inputs for the reference CPU solver, and its timing.

TEST INFRASTRUCTURE ONLY: used by bench.py's cpu_baseline leg (and tests) to
time the reference's own single-threaded CPU solver,
hydro-cal-fortan/calculate.cpp, built unmodified from the reference checkout
into oracle/_ref/hydro_cc by ``make -C oracle ref``.  The product never uses
it.

That solver reads a fixed directory, ``./origin/changjiangkou/`` under its
working directory (hydro-cal-fortan/read_data.hh:39, calculate.cpp:10), in
the text formats of its loaders (calculate.cpp:749-939): TIME, GIRD, DEPTH,
BOUNDARY, CALTIME (one value per line), SOURCES/PNAC, PNAP, PKLAS (a header
line, then ``NO v0 v1 v2 v3`` per cell), PZBC, INITIALLEVEL, INITIALU1,
INITIALV1, CV (a header line, then one value per cell), PXY (a header line,
then ``NO X Y`` per node), MBZ, MBQ (a count line, then one line per
boundary series).  Its DT is an integer (calculate.cpp:20), and its
time-series paths index past their arrays (ZT/QT transposed,
calculate.cpp:1052 vs 276/320/1170; QZSTIME2 never sized,
calculate.cpp:1094), so the case is written without series: NZ = NQ = 0,
discharge (KLAS 10) and level (KLAS 1) edges written as walls (KLAS 4),
NDAYS = 1 and MDT = steps * DT, DT = 1 s.  It then runs ``steps`` timesteps
of the same mesh, bed, Manning field and initial state.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
CC_PATH = os.path.join(_HERE, "_ref", "hydro_cc")


def _lines(path, header, rows, fmt):
    with open(path, "w") as f:
        f.write(header + "\n")
        np.savetxt(f, rows, fmt=fmt)


def write_case(host, root, steps, dt=1):
    """Write ``host`` (a HostMesh) as the solver's input directory under root."""
    d = os.path.join(root, "origin", "changjiangkou")
    src = os.path.join(d, "SOURCES")
    os.makedirs(src, exist_ok=True)
    C = host.CELL
    a = host.arrays
    XP, YP = a["XP"], a["YP"]
    NOD = len(XP)

    def scalars(name, *vals):
        with open(os.path.join(d, name + ".DAT"), "w") as f:
            for v in vals:
                f.write(f"{v}\n")

    scalars("TIME", int(steps * dt), 1)                 # MDT, NDAYS
    scalars("GIRD", NOD, C)
    scalars("DEPTH", repr(float(host.HM1)), repr(float(host.HM2)))
    scalars("BOUNDARY", 0, 0, 0, 0, 0, 0)               # NZ, NQ, -, NHQ, -, NDI
    scalars("CALTIME", 0, int(dt))
    no = np.arange(1, C + 1, dtype=np.int64)[:, None]
    nap = a["NAP"].reshape(C, 4).astype(np.int64)
    nac = a["NAC"].reshape(C, 4).astype(np.int64)
    klas = a["KLAS"].reshape(C, 4).astype(np.int64)
    klas = np.where((klas == 1) | (klas == 10), 4, klas)
    _lines(os.path.join(src, "PNAC.DAT"), "NO NAC", np.hstack([no, nac]), "%d")
    _lines(os.path.join(src, "PNAP.DAT"), "NO NAP", np.hstack([no, nap]), "%d")
    _lines(os.path.join(src, "PKLAS.DAT"), "NO KLAS", np.hstack([no, klas]), "%d")
    _lines(os.path.join(src, "PZBC.DAT"), "ZBC", a["ZBC"].astype(np.float64)[:, None], "%.17g")
    xy = np.column_stack([np.arange(1, NOD + 1), XP.astype(np.float64), YP.astype(np.float64)])
    _lines(os.path.join(src, "PXY.DAT"), "NO X Y", xy, ["%d", "%.17g", "%.17g"])
    for name in ("MBZ", "MBQ"):
        with open(os.path.join(src, name + ".DAT"), "w") as f:
            f.write("0\n")
    _lines(os.path.join(d, "INITIALLEVEL.DAT"), "Z", a["Z"].astype(np.float64)[:, None], "%.17g")
    _lines(os.path.join(d, "INITIALU1.DAT"), "U", a["U"].astype(np.float64)[:, None], "%.17g")
    _lines(os.path.join(d, "INITIALV1.DAT"), "V", a["V"].astype(np.float64)[:, None], "%.17g")
    n = np.sqrt(a["FNC"].astype(np.float64) / 9.81)
    _lines(os.path.join(d, "CV.DAT"), "CV", n[:, None], "%.17g")
    return d


def run_case(root, exe=CC_PATH, timeout=600):
    """Run the solver in ``root``; wall seconds of the whole process."""
    t0 = time.perf_counter()
    subprocess.run([exe], cwd=root, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                   timeout=timeout)
    return time.perf_counter() - t0


def time_reference(host, steps_short, steps_long, exe=CC_PATH):
    """Seconds per timestep of the reference CPU solver on ``host``: two runs
    of steps_short and steps_long steps, so reading the input and the
    geometry preprocessing cancel out.  Returns (seconds_per_step, detail)."""
    if not os.path.exists(exe):
        raise FileNotFoundError(exe)
    out = {}
    for steps in (steps_short, steps_long):
        root = tempfile.mkdtemp(prefix="hydro_cc_")
        try:
            write_case(host, root, steps)
            out[steps] = run_case(root, exe)
        finally:
            shutil.rmtree(root, ignore_errors=True)
    per_step = (out[steps_long] - out[steps_short]) / (steps_long - steps_short)
    return per_step, out
