"""ctypes wrapper of the CPU oracle (oracle/_build/libhydro_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never uses it.

``OracleRun`` owns host copies of a mesh's mutable arrays (state, KLAS,
FLUX0..3) and steps them with the restatement in hydro_oracle.cpp.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_int32, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libhydro_oracle.so")


class OracleMesh(ctypes.Structure):
    _fields_ = [
        ("dtype", c_int32), ("variant", c_int32), ("CELL", c_int32), ("NDAYS", c_int32), ("NHQ", c_int32),
        ("MDT", c_int32), ("DT", c_double), ("HM1", c_double), ("HM2", c_double), ("JL", c_double),
        ("H", c_void_p), ("U", c_void_p), ("V", c_void_p), ("Z", c_void_p), ("W", c_void_p),
        ("ZBC", c_void_p), ("ZB1", c_void_p), ("AREA", c_void_p), ("FNC", c_void_p),
        ("BoundaryFeature", c_void_p), ("NHQ1", c_void_p),
        ("ZW", c_void_p), ("QW", c_void_p), ("QT", c_void_p), ("DQT", c_void_p), ("ZT", c_void_p),
        ("DZT", c_void_p), ("NAC", c_void_p), ("KLAS", c_void_p),
        ("SIDE", c_void_p), ("COSF", c_void_p), ("SINF", c_void_p), ("SLCOS", c_void_p), ("SLSIN", c_void_p),
        ("FLUX0", c_void_p), ("FLUX1", c_void_p), ("FLUX2", c_void_p), ("FLUX3", c_void_p),
    ]


_lib = None


def build():
    """Compile the oracle with its Makefile (g++ only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _bind(lib):
    lib.oracle_calculate_flux.argtypes = [POINTER(OracleMesh), c_int, c_int]
    lib.oracle_update_cell.argtypes = [POINTER(OracleMesh)]
    lib.oracle_step_n.argtypes = [POINTER(OracleMesh), c_int, POINTER(c_int), POINTER(c_int)]
    lib.oracle_step_n.restype = c_int
    lib.oracle_osher.argtypes = [c_int, POINTER(c_double), POINTER(c_double), c_double, c_double,
                                 POINTER(c_double)]
    lib.oracle_precalculate.argtypes = [c_int, c_int] + [c_void_p] * 6 + [c_double] + [c_void_p] * 9
    lib.oracle_set_threads.argtypes = [c_int]
    lib.oracle_set_threads.restype = c_int
    return lib


def load_mutant(k, outdir):
    """The oracle built with planted mistake k (ORACLE_MUTANT, for the
    mutation test of the pins), compiled into `outdir`."""
    out = os.path.join(outdir, f"libhydro_oracle_mutant{k}.so")
    subprocess.run(["make", "-s", "-C", _HERE, "mutant", f"K={k}", f"OUT={out}"], check=True)
    return _bind(ctypes.CDLL(out))


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        lib.oracle_calculate_flux.argtypes = [POINTER(OracleMesh), c_int, c_int]
        lib.oracle_update_cell.argtypes = [POINTER(OracleMesh)]
        lib.oracle_step_n.argtypes = [POINTER(OracleMesh), c_int, POINTER(c_int), POINTER(c_int)]
        lib.oracle_step_n.restype = c_int
        lib.oracle_osher.argtypes = [c_int, POINTER(c_double), POINTER(c_double), c_double, c_double,
                                     POINTER(c_double)]
        lib.oracle_precalculate.argtypes = [c_int, c_int] + [c_void_p] * 6 + [c_double] + [c_void_p] * 9
        lib.oracle_set_threads.argtypes = [c_int]
        lib.oracle_set_threads.restype = c_int
        _lib = lib
    return _lib


def set_threads(n):
    """Threads of the oracle's slot and cell loops (OpenMP; 1 by default,
    the same bits at any count).  Returns the previous count."""
    return load().oracle_set_threads(int(n))


MUTABLE = ("H", "U", "V", "Z", "W", "KLAS", "FLUX0", "FLUX1", "FLUX2", "FLUX3")


class OracleRun:
    """CPU run of one mesh; `variant` 0 = A, 1 = B."""

    def __init__(self, mesh, variant: int = 0, lib=None):
        self.lib = lib if lib is not None else load()
        self.mesh = mesh
        self.a = {k: (v.copy() if k in MUTABLE else v) for k, v in mesh.arrays.items()}
        om = OracleMesh()
        om.dtype = 0 if mesh.dtype == np.float64 else 1
        om.variant = variant
        om.CELL, om.NDAYS, om.NHQ, om.MDT = mesh.CELL, mesh.NDAYS, mesh.NHQ, mesh.MDT
        om.DT, om.HM1, om.HM2, om.JL = mesh.DT, mesh.HM1, mesh.HM2, mesh.JL
        for name, _ in om._fields_:
            if name in self.a:
                setattr(om, name, self.a[name].ctypes.data_as(c_void_p).value)
        self.om = om
        self.day = c_int(0)
        self.kt = c_int(1)

    def calculate_flux(self, kt, jt):
        self.lib.oracle_calculate_flux(ctypes.byref(self.om), kt, jt)

    def update_cell(self):
        self.lib.oracle_update_cell(ctypes.byref(self.om))

    def step(self, n=1):
        done = self.lib.oracle_step_n(ctypes.byref(self.om), int(n), ctypes.byref(self.day), ctypes.byref(self.kt))
        if done != n:
            raise RuntimeError(f"oracle ran {done} of {n} steps (NDAYS cap)")

    def __getitem__(self, k):
        return self.a[k]


def osher(QL, QR, FIL, H_pre, dtype=0):
    lib = load()
    ql = (c_double * 3)(*QL)
    qr = (c_double * 3)(*QR)
    out = (c_double * 4)()
    lib.oracle_osher(dtype, ql, qr, float(FIL), float(H_pre), out)
    return np.array(out[:])


def precalculate(CELL, NOD, NAP, XP, YP, ZBC, Z, FNC0, HM1):
    """fp64 restatement of MeshData::preCalculate; returns a dict of outputs."""
    lib = load()
    NAP = np.ascontiguousarray(NAP, dtype=np.int32)
    XP, YP, ZBC, Z, FNC0 = (np.array(a, dtype=np.float64) for a in (XP, YP, ZBC, Z, FNC0))
    out = {k: np.zeros(4 * CELL) for k in ("SIDE", "SINF", "COSF", "SLCOS", "SLSIN")}
    out.update({k: np.zeros(CELL) for k in ("AREA", "ZB1", "H", "FNC")})
    p = lambda a: a.ctypes.data_as(c_void_p)
    lib.oracle_precalculate(CELL, NOD, p(NAP), p(XP), p(YP), p(ZBC), p(Z), p(FNC0), float(HM1),
                            p(out["SIDE"]), p(out["SINF"]), p(out["COSF"]), p(out["SLCOS"]), p(out["SLSIN"]),
                            p(out["AREA"]), p(out["ZB1"]), p(out["H"]), p(out["FNC"]))
    out["Z"] = Z
    out["XP"], out["YP"] = XP, YP
    return out
