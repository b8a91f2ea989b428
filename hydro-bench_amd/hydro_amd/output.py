"""The reference's text OUTPUT files, written from a state.

``MeshData::outputToFile(jt, kt)`` (cuda_native_impl/hydro-cal-src/src/
mesh.cpp:535-839) appends one frame to four files and writes the geometry
file once:

  SIDE.OUT     at (jt 0, kt 1) only: COSF, SINF, SIDE per cell (4 values)
               and AREA, in the stream's default format (%g, 6 digits);
  H2U2V2.OUT   a header line, then H2=, U2=, V2= blocks;
  ZUV.OUT      the same header, then Z2=, W2=, FI= blocks;
  TIMELOG.OUT  jt2 / NDAYS, fixed with 4 decimals;
  XY-TEC.DAT   a Tecplot zone: node X, Y in the original frame (the
               preprocessing's shift added back), cell H2, Z2, U2, V2, W2,
               then the cell connectivity NAP.

Blocks print each value as ``setw(10) fixed setprecision(4)``; before the
first value and every tenth a newline and five spaces, and another at every
hundredth.  Dry cells (H <= HM1) print H2 = U2 = V2 = W2 = FI = 0 and Z2 =
ZBC; values below 1e-9 in magnitude print as +0.  W2 is sqrt(U^2 + V^2)
rounded to float; FI is the flow angle in degrees (MeshData::FI,
mesh.cpp:841-878: pi taken as 3.1416, the result scaled by 57.298).  jt2 is
jt, or jt + 1 when kt != 1.

The C++ streams keep ``fixed`` and the last precision from one insertion to
the next, so a header's ``DT=`` field prints in the default format in the
first frame of a file (``  4``, ``0.5``) and as fixed with 4 decimals in
every later one (``4.0000``, ``0.5000``).

Arithmetic (W2, FI, jt2 / NDAYS, XP + XIMIN) runs in the mesh's precision,
as the reference does in its Real type.

:func:`run_with_output` drives a device mesh the way the reference
benchmark's ``runStepsWithOutput`` does (cuda_native_impl/benchmark.cu:
98-124): a frame at (0, 1), then one after the last step of every
``ntoutput``-th day and after the last step.

:func:`compare_dirs` compares two OUTPUT directories the way the
reference's scripts/compare_output_files.py reports them: per file whether
the bytes match and how many lines differ, and per frame and block (H2, U2,
V2, Z2, W2, FI) the maximum, mean and 99th percentile of the absolute
differences and how many values differ.  ``python -m hydro_amd.output A B``
prints that report as JSON.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

FILES = ("SIDE.OUT", "H2U2V2.OUT", "ZUV.OUT", "TIMELOG.OUT", "XY-TEC.DAT")


def _snap(v):
    v = np.array(v, copy=True)
    v[np.abs(v) < 1e-9] = 0
    return v


def flow_angle(x, y):
    """MeshData::FI over arrays, in degrees (x 57.298), in the arrays'
    precision.  Components below 1e-9 count as zero; when one is zero the
    reference's four tests run in sequence and the later wins, so
    FI(0, 0) = 0."""
    x = _snap(np.asarray(x))
    y = _snap(np.asarray(y))
    R = x.dtype.type
    pi = R(3.1416)
    prod = x * y
    w = np.arctan2(np.abs(y), np.abs(x))
    fi = np.where(prod > 0, np.where(x > 0, w, pi + w), np.where(y > 0, pi - w, R(2) * pi - w))
    zero = prod == 0
    fi = np.where(zero & (x == 0) & (y >= 0), pi / R(2), fi)
    fi = np.where(zero & (x == 0) & (y < 0), R(3) * pi / R(2), fi)
    fi = np.where(zero & (y == 0) & (x >= 0), R(0), fi)
    fi = np.where(zero & (y == 0) & (x < 0), pi, fi)
    return (fi * R(57.298)).astype(x.dtype)


def _block(label, values, first):
    out = ["     \n     " if first else "\n     \n     ", label]
    vals = np.char.mod("%10.4f", np.asarray(values, dtype=np.float64))
    for i, s in enumerate(vals):
        if i % 10 == 0:
            out.append("\n     ")
        if i != 0 and i % 100 == 0:
            out.append("\n     ")
        out.append(s)
    return "".join(out)


def _lines(values, lead):
    """Values as ``%.4f `` with a newline before every tenth (and before the
    first when `lead`)."""
    s = np.char.mod("%.4f ", np.asarray(values, dtype=np.float64))
    parts = []
    for i, x in enumerate(s):
        if i % 10 == 0 and (lead or i != 0):
            parts.append("\n")
        parts.append(x)
    return "".join(parts)


class OutputFiles:
    """The five OUTPUT files of one mesh in `out_dir` (truncated on open);
    :meth:`frame` appends one outputToFile frame.

    `host` is a :class:`~hydro_amd.mesh.HostMesh` (or any object with
    ``arrays``, ``CELL``, ``HM1``, ``DT``, ``NDAYS``); its arrays give COSF,
    SINF, SIDE, AREA, ZBC, NAP and the shifted node coordinates XP, YP with
    the shift XYMIN.  `dt` and `ndays` replace the mesh's DT and NDAYS in
    the header and the time log."""

    def __init__(self, out_dir, host, dt=None, ndays=None):
        os.makedirs(out_dir, exist_ok=True)
        a = host.arrays
        self.a = a
        self.C = int(host.CELL)
        self.R = np.asarray(a["AREA"]).dtype.type
        self.HM1 = float(host.HM1)
        self.DT = self.R(host.DT if dt is None else dt)
        self.NDAYS = int(host.NDAYS if ndays is None else ndays)
        self.xymin = np.asarray(a.get("XYMIN", np.zeros(2))).astype(self.R)
        self.f = {n: open(os.path.join(out_dir, n), "w") for n in FILES}
        self.fixed = {"H2U2V2.OUT": False, "ZUV.OUT": False}

    def close(self):
        for f in self.f.values():
            f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _side(self):
        a, C = self.a, self.C
        lines = []
        for name in ("COSF", "SINF", "SIDE"):
            v = np.asarray(a[name], dtype=np.float64).reshape(C, 4)
            lines.append(name + "\n")
            for i in range(C):
                lines.append(f"{i + 1}    " + "".join("%.6g     " % x for x in v[i]) + "\n")
        lines.append("AREA\n")
        for i, x in enumerate(np.asarray(a["AREA"], dtype=np.float64)):
            lines.append(f"{i + 1}    {x:.6g}\n")
        self.f["SIDE.OUT"].write("".join(lines))

    def _header(self, name, jt2, kt):
        if self.fixed[name]:
            dt = "%.4f" % float(self.DT)
        else:
            dt = ("%g" % float(self.DT)).rjust(3)
        return f" \n \n JT={jt2:5d}  KT={kt - 1:5d}  DT={dt} SEC  T= 0H  NSF= 0/0  WEC=0/  CQL=0  INE=0\n"

    def frame(self, jt, kt, state):
        """Append the frame of `state` (H, U, V, Z arrays of CELL entries in
        the mesh's cell order) at day `jt`, step `kt`."""
        C, R = self.C, self.R
        H = np.asarray(state["H"])[:C].astype(R)
        U = np.asarray(state["U"])[:C].astype(R)
        V = np.asarray(state["V"])[:C].astype(R)
        Z = np.asarray(state["Z"])[:C].astype(R)
        wet = H > self.HM1
        zero = R(0)
        h2 = _snap(np.where(wet, H, zero))
        u2 = _snap(np.where(wet, U, zero))
        v2 = _snap(np.where(wet, V, zero))
        z2 = _snap(np.where(wet, Z, np.asarray(self.a["ZBC"]).astype(R)))
        w2 = np.where(wet, np.sqrt(U * U + V * V), zero).astype(np.float32)
        w2[np.abs(w2) < np.float32(1e-9)] = 0
        fi = np.where(wet, flow_angle(U, V), zero)
        jt2 = jt if kt == 1 else jt + 1
        if jt == 0 and kt == 1:
            self._side()
        self.f["H2U2V2.OUT"].write(self._header("H2U2V2.OUT", jt2, kt) + _block("H2=", h2, True)
                                   + _block("U2=", u2, False) + _block("V2=", v2, False) + " \n")
        self.fixed["H2U2V2.OUT"] = True
        self.f["ZUV.OUT"].write(self._header("ZUV.OUT", jt2, kt) + _block("Z2=", z2, True)
                                + _block("W2=", w2, False) + _block("FI=", fi, False) + " \n")
        self.fixed["ZUV.OUT"] = True
        self.f["TIMELOG.OUT"].write("%.4f\n" % float(R(jt2) / R(self.NDAYS)))
        self._tec(h2, z2, u2, v2, w2)
        for fh in self.f.values():
            fh.flush()

    def _tec(self, h2, z2, u2, v2, w2):
        a = self.a
        XP = np.asarray(a["XP"]).astype(self.R) + self.xymin[0]
        YP = np.asarray(a["YP"]).astype(self.R) + self.xymin[1]
        out = [' VARIABLES = "X", "Y", "H2", "Z2","U2","V2","W2"\n',
               f"ZONE N={len(XP)}, E={self.C}, DATAPACKING=BLOCK, ZONETYPE=FEQUADRILATERAL\n",
               "VARLOCATION=([3-7]=CELLCENTERED)\n",
               _lines(XP, False) + "\n", _lines(YP, False) + "\n",
               _lines(h2, True)]                   # the H2 block ends without a newline
        for v in (z2, u2, v2, w2):
            out.append(_lines(v, False) + "\n")
        nap = np.asarray(a["NAP"]).reshape(self.C, 4)
        out.append("".join("%d %d %d %d \n" % tuple(r) for r in nap.tolist()))
        self.f["XY-TEC.DAT"].write("".join(out))


def run_with_output(dev, out_dir, steps, ntoutput=1, mode="fused"):
    """Run `steps` timesteps of a DeviceMesh from its current state (at the
    start of its time grid) and write the OUTPUT files: a frame at (0, 1),
    then after the last step of every day whose index + 1 is a multiple of
    `ntoutput`, and after the last step.  Steps between frames run as one
    batch.  Returns the list of (jt, kt) frames written."""
    host = dev.host
    spd = host.steps_per_day
    if steps > host.max_steps:
        raise ValueError(f"{steps} steps exceed the NDAYS cap {host.max_steps}")
    ntoutput = max(int(ntoutput), 1)
    frames = []
    with OutputFiles(out_dir, host) as out:
        out.frame(0, 1, dev.fetch("H", "U", "V", "Z"))
        frames.append((0, 1))
        s = pending = 0
        for day in range(host.NDAYS):
            for kt in range(1, spd):
                if s >= steps:
                    break
                s += 1
                pending += 1
                last = kt + 1 == spd or s == steps
                if last and ((day + 1) % ntoutput == 0 or s == steps):
                    dev.step(pending, mode)
                    pending = 0
                    out.frame(day, kt + 1, dev.fetch("H", "U", "V", "Z"))
                    frames.append((day, kt + 1))
        if pending:
            dev.step(pending, mode)
    return frames


_LABELS = ("H2=", "U2=", "V2=", "Z2=", "W2=", "FI=")


def read_blocks(path):
    """{(frame index, label): values} of an H2U2V2.OUT or ZUV.OUT file."""
    out = {}
    frame, label, vals = -1, None, []

    def close():
        if label is not None:
            out[(frame, label[:2])] = np.array(vals, dtype=np.float64)

    with open(path) as f:
        for line in f:
            t = line.strip()
            if t.startswith("JT="):
                close()
                frame, label, vals = frame + 1, None, []
            elif t in _LABELS:
                close()
                label, vals = t, []
            elif label is not None and t:
                vals.extend(float(x) for x in t.split())
    close()
    return out


def compare_dirs(a, b):
    """Report of OUTPUT directory `b` against `a` (see the module notes)."""
    rep = {"files": {}, "blocks": {}}
    for fn in FILES:
        pa, pb = os.path.join(a, fn), os.path.join(b, fn)
        if not (os.path.isfile(pa) and os.path.isfile(pb)):
            rep["files"][fn] = {"present": [os.path.isfile(pa), os.path.isfile(pb)]}
            continue
        with open(pa, "rb") as f:
            da = f.read()
        with open(pb, "rb") as f:
            db = f.read()
        la, lb = da.split(b"\n"), db.split(b"\n")
        differ = sum(x != y for x, y in zip(la, lb)) + abs(len(la) - len(lb))
        rep["files"][fn] = {"byte_equal": da == db, "lines": [len(la), len(lb)], "lines_differ": differ}
        if fn in ("H2U2V2.OUT", "ZUV.OUT"):
            ba, bb = read_blocks(pa), read_blocks(pb)
            for key in sorted(set(ba) & set(bb)):
                x, y = ba[key], bb[key]
                if len(x) != len(y):
                    rep["blocks"][f"{key[0]}/{key[1]}"] = {"lengths": [len(x), len(y)]}
                    continue
                d = np.abs(x - y)
                rep["blocks"][f"{key[0]}/{key[1]}"] = {
                    "max_abs": float(d.max()) if d.size else 0.0,
                    "mean_abs": float(d.mean()) if d.size else 0.0,
                    "p99": float(np.percentile(d, 99)) if d.size else 0.0,
                    "n_differ": int(np.count_nonzero(d)), "n": int(d.size)}
    rep["all_byte_equal"] = all(v.get("byte_equal", False) for v in rep["files"].values())
    return rep


def main(argv=None):
    import argparse

    ap = argparse.ArgumentParser(description="compare two OUTPUT directories")
    ap.add_argument("a")
    ap.add_argument("b")
    ap.add_argument("--out", default=None, help="also write the report here")
    args = ap.parse_args(argv)
    rep = compare_dirs(args.a, args.b)
    text = json.dumps(rep, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    print(text)
    return 0 if rep["all_byte_equal"] else 1


if __name__ == "__main__":
    sys.exit(main())
