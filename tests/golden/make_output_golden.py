"""Writes tests/golden/output_frames.npz: OUTPUT files (SIDE.OUT, H2U2V2.OUT,
ZUV.OUT, TIMELOG.OUT, XY-TEC.DAT) as the reference's Python writer
(taichi_impl/output_writer.py, its restatement of MeshData::outputToFile,
cuda_native_impl/hydro-cal-src/src/mesh.cpp:535-839) produces them for a
small mesh and three seeded states, with the inputs needed to write them
again.

The reference's C++ writer cannot be built here (mesh.cpp includes
common.hpp, which includes cuda_runtime.h), and the reference holds no
OUTPUT files in this format, so these vectors are what pins the writer.

The states cover dry cells (H <= HM1, H == HM1), exact zeros, values below
the 1e-9 snap in both signs, values that print as -0.0000, every quadrant
and axis of the flow angle, and bed levels near zero.  Two header cases:
DT = 4 (integer) and DT = 0.5.

Run from the repository root with the reference checkout at /root/reference.
"""
import importlib.util
import os
import tempfile
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "hydro-bench_amd"))
WRITER = "/root/reference/taichi_impl/output_writer.py"
FILES = ("SIDE.OUT", "H2U2V2.OUT", "ZUV.OUT", "TIMELOG.OUT", "XY-TEC.DAT")
CASES = (("dt4", 4.0, 3), ("dt05", 0.5, 7))          # name, DT, NDAYS


def states(host, rng, n):
    C = host.CELL
    zbc = host.arrays["ZBC"].astype(np.float64).copy()
    out = []
    for f in range(n):
        H = rng.uniform(0.0, 2.5, C)
        dry = rng.random(C) < 0.15
        H[dry] = rng.uniform(0.0, host.HM1, dry.sum())
        H[rng.random(C) < 0.03] = host.HM1
        U = rng.normal(0.0, 0.8, C)
        V = rng.normal(0.0, 0.8, C)
        for A in (U, V):
            pick = rng.random(C)
            A[pick < 0.08] = 0.0
            A[(pick >= 0.08) & (pick < 0.12)] = 3e-12
            A[(pick >= 0.12) & (pick < 0.16)] = -7e-10
            A[(pick >= 0.16) & (pick < 0.19)] = -2e-5          # prints -0.0000
            A[(pick >= 0.19) & (pick < 0.21)] = 1e-9            # kept, not snapped
        Z = zbc + H
        near = rng.random(C) < 0.05
        Z[near] = rng.choice([0.0, 4e-10, -3e-10, -0.25, -1e-5], near.sum())
        out.append({"H": H, "U": U, "V": V, "Z": Z})
    return out


def main(out=os.path.join(ROOT, "tests", "golden", "output_frames.npz")):
    from hydro_amd.mesh import make_channel_mesh

    spec = importlib.util.spec_from_file_location("reference_output_writer", WRITER)
    ref = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ref)

    host = make_channel_mesh(seed=4242, cells=230, DT=4.0, MDT=60, steps=30, rotate=0.4)
    # zero bed on a few cells, so that dry Z2 = ZBC snaps
    host.arrays["ZBC"][[3, 57, 111]] = [0.0, 2e-10, -5e-10]
    a = host.arrays
    C = host.CELL
    nap = np.zeros((5, C + 1), dtype=np.int64)
    nap[1:, 1:] = a["NAP"].reshape(C, 4).T
    mesh = {"CELL": C, "KLAS": a["KLAS"], "COSF": a["COSF"], "SINF": a["SINF"], "SIDE": a["SIDE"],
            "AREA": a["AREA"], "ZBC": a["ZBC"], "NAP": nap,
            "XP": np.concatenate([[0.0], a["XP"]]), "YP": np.concatenate([[0.0], a["YP"]]),
            "XIMIN": float(a["XYMIN"][0]), "YIMIN": float(a["XYMIN"][1]), "HM1": host.HM1}
    rng = np.random.default_rng(99)
    frames = [(0, 1), (0, 15), (1, 15)]
    st = states(host, rng, len(frames))
    keep = {k: a[k] for k in ("COSF", "SINF", "SIDE", "AREA", "ZBC", "NAP", "XP", "YP", "XYMIN")}
    keep.update(CELL=C, HM1=host.HM1, frames=np.array(frames))
    for i, s in enumerate(st):
        for k, v in s.items():
            keep[f"state{i}_{k}"] = v
    tmpdir = tempfile.TemporaryDirectory()
    tmp = tmpdir.name
    for name, dt, ndays in CASES:
        d = os.path.join(tmp, name)
        with ref.OutputWriter(d, mesh, dt, ndays) as w:
            for (jt, kt), s in zip(frames, st):
                w.write_frame(jt, kt, s)
        keep[f"{name}_DT"] = dt
        keep[f"{name}_NDAYS"] = ndays
        for fn in FILES:
            with open(os.path.join(d, fn), "rb") as f:
                keep[f"{name}_{fn}"] = np.frombuffer(f.read(), dtype=np.uint8)
    tmpdir.cleanup()
    np.savez_compressed(out, **keep)
    print(out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
