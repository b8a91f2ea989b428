"""Alignment checker: two states of the same mesh, field statistics,
conservation sums and a PASS / FAIL verdict.

The statistics and the verdict gates follow the reference's alignment
checker (scripts/check_correctness.py:519-807, thresholds of
scripts/ALIGNMENT_PLAN.md): per field the maximum and mean absolute
difference, its 50/90/99/99.9th percentiles, how many entries differ by more
than 1e-13/1e-11/1e-9/1e-7, the fraction that is bit-identical and the worst
cells with their KLAS codes; per state the mass, momentum, kinetic and
potential energy and the KLAS 10 / KLAS 1 boundary inflow (F0 * SIDE over
those slots), summed exactly (``math.fsum``, where the reference uses
Neumaier summation), and their relative differences.

Gates:
  fp64, any step       every state field max_abs < 1e-9 and p99 < 1e-11,
                       every conservation sum within 1e-12 relative;
  fp32, step 1         every state field >= 99% bit-identical and
                       max_abs < 1e-6, fluxes max_abs < 1e-5;
  fp32, later steps    every state field p99 < 1e-3, conservation within
                       1e-5 relative;
and never a NaN or Inf in the state.

States come from :func:`read_dump` (the binary dump of ``hydro_dump``, the
reference benchmark's format: int32 CELL; H U V Z [CELL]; SLCOS SLSIN SIDE
FLUX0..3 [4*CELL]) or from dicts of arrays.  ``python -m hydro_amd.align
REF GOT --config NAME --precision fp64 --step N`` compares two dumps of one
of the seeded workloads and prints the report as JSON.
"""
from __future__ import annotations

import json
import math
import sys

import numpy as np

GRAVITY = 9.81
TOLERANCES = {
    "fp64_state_max_abs": 1e-9,
    "fp64_state_p99": 1e-11,
    "fp64_conservation_rel": 1e-12,
    "fp32_step1_bit_exact_frac": 0.99,
    "fp32_step1_state_max_abs": 1e-6,
    "fp32_step1_flux_max_abs": 1e-5,
    "fp32_long_p99": 1e-3,
    "fp32_long_conservation_rel": 1e-5,
}
STATE = ("H", "U", "V", "Z", "W")
FLUX = ("F0", "F1", "F2", "F3")


def read_dump(path, dtype=np.float64):
    """A dump written by hydro_dump (or write_dump)."""
    with open(path, "rb") as f:
        C = int(np.frombuffer(f.read(4), dtype=np.int32)[0])
        out = {}
        for k in ("H", "U", "V", "Z"):
            out[k] = np.frombuffer(f.read(C * np.dtype(dtype).itemsize), dtype=dtype).copy()
        for k in ("SLCOS", "SLSIN", "SIDE", "F0", "F1", "F2", "F3"):
            out[k] = np.frombuffer(f.read(4 * C * np.dtype(dtype).itemsize), dtype=dtype).copy()
    out["W"] = np.sqrt(out["U"] * out["U"] + out["V"] * out["V"])
    return out


def write_dump(path, state, dtype=np.float64):
    """The same format from arrays (H U V Z, SLCOS SLSIN SIDE, FLUX0..3 or F0..F3)."""
    C = len(state["H"])
    with open(path, "wb") as f:
        f.write(np.int32(C).tobytes())
        for k in ("H", "U", "V", "Z"):
            f.write(np.ascontiguousarray(state[k], dtype=dtype).tobytes())
        for k in ("SLCOS", "SLSIN", "SIDE"):
            f.write(np.ascontiguousarray(state[k], dtype=dtype).tobytes())
        for i in range(4):
            a = state.get(f"F{i}", state.get(f"FLUX{i}"))
            f.write(np.ascontiguousarray(a, dtype=dtype).tobytes())


def _percentiles(d):
    if d.size == 0:
        return {p: 0.0 for p in (50, 90, 99, 99.9)}
    return {p: float(np.percentile(d, p)) for p in (50, 90, 99, 99.9)}


def field_stats(name, ref, got, klas=None, nac=None, top=3):
    a = np.asarray(ref, dtype=np.float64)
    b = np.asarray(got, dtype=np.float64)
    n = min(len(a), len(b))
    a, b = a[:n], b[:n]
    fin = np.isfinite(a) & np.isfinite(b)
    if not fin.any():
        return {"all_nonfinite": True, "n_total": int(n)}
    d = np.abs(a[fin] - b[fin])
    out = {
        "max_abs": float(d.max()),
        "mean_abs": float(d.mean()),
        "percentiles": _percentiles(d),
        "threshold_counts": {f"diff_gt_{t:.0e}": int((d > t).sum()) for t in (1e-13, 1e-11, 1e-9, 1e-7)},
        "bit_exact_frac": float(np.count_nonzero(a[fin] == b[fin])) / float(d.size),
        "n_finite": int(fin.sum()),
        "n_total": int(n),
    }
    full = np.where(fin, np.abs(a - b), -1.0)
    worst = []
    slot = name in FLUX
    for i in np.argsort(full)[::-1][:top]:
        if full[i] < 0:
            continue
        rec = {"idx": int(i), "ref": float(a[i]), "got": float(b[i]), "abs_diff": float(full[i])}
        if klas is not None:
            cell = int(i // 4) if slot else int(i)
            rec["cell"] = cell
            rec["klas"] = int(klas[i]) if slot else [int(klas[4 * cell + j]) for j in range(4)]
            if nac is not None:
                rec["neighbours"] = [int(nac[4 * cell + j]) for j in range(4)]
        worst.append(rec)
    out["worst_cells"] = worst
    return out


def conservation(state, area, klas, side):
    H = np.asarray(state["H"], dtype=np.float64)
    U = np.asarray(state["U"], dtype=np.float64)
    V = np.asarray(state["V"], dtype=np.float64)
    a = np.asarray(area, dtype=np.float64)
    F0 = np.asarray(state["F0"], dtype=np.float64)
    s = np.asarray(side, dtype=np.float64)
    k = np.asarray(klas)
    return {
        "mass": math.fsum(H * a),
        "momentum_x": math.fsum(H * U * a),
        "momentum_y": math.fsum(H * V * a),
        "kinetic_energy": math.fsum(0.5 * H * (U * U + V * V) * a),
        "potential_energy": math.fsum(0.5 * GRAVITY * H * H * a),
        "klas10_inflow": math.fsum(F0 * s * (k == 10)),
        "klas1_inflow": math.fsum(F0 * s * (k == 1)),
    }


def conservation_diffs(ref, got):
    out = {}
    for k in ref:
        r, g = ref[k], got[k]
        out[k] = {"ref": r, "got": g, "abs_diff": abs(r - g), "rel_diff": abs(r - g) / max(abs(r), 1e-30)}
    return out


def verdict(precision, step, stats, cons, nonfinite):
    if nonfinite:
        return "FAIL", f"NaN/Inf in the state: {nonfinite}"
    if any(stats[f].get("all_nonfinite") for f in STATE):
        return "FAIL", "non-finite state field"
    T = TOLERANCES
    if precision == "fp64":
        for f in STATE:
            if stats[f]["max_abs"] >= T["fp64_state_max_abs"]:
                return "FAIL", f"{f}.max_abs={stats[f]['max_abs']:.3e} >= 1e-9"
            if stats[f]["percentiles"][99] >= T["fp64_state_p99"]:
                return "FAIL", f"{f}.p99={stats[f]['percentiles'][99]:.3e} >= 1e-11"
        for k, v in cons.items():
            if v["rel_diff"] >= T["fp64_conservation_rel"]:
                return "FAIL", f"conservation/{k} rel={v['rel_diff']:.3e} >= 1e-12"
        return "PASS", "fp64 thresholds met"
    if step == 1:
        for f in STATE:
            if stats[f]["bit_exact_frac"] < T["fp32_step1_bit_exact_frac"]:
                return "FAIL", f"{f}.bit_exact_frac={stats[f]['bit_exact_frac']:.4f} < 0.99"
            if stats[f]["max_abs"] >= T["fp32_step1_state_max_abs"]:
                return "FAIL", f"{f}.max_abs={stats[f]['max_abs']:.3e} >= 1e-6"
        for f in FLUX:
            s = stats.get(f)
            if s and not s.get("all_nonfinite") and s["max_abs"] >= T["fp32_step1_flux_max_abs"]:
                return "FAIL", f"{f}.max_abs={s['max_abs']:.3e} >= 1e-5"
        return "PASS", "fp32 step=1 thresholds met"
    for f in STATE:
        if stats[f]["percentiles"][99] >= T["fp32_long_p99"]:
            return "FAIL", f"{f}.p99={stats[f]['percentiles'][99]:.3e} >= 1e-3"
    for k, v in cons.items():
        if v["rel_diff"] >= T["fp32_long_conservation_rel"]:
            return "FAIL", f"conservation/{k} rel={v['rel_diff']:.3e} >= 1e-5"
    return "PASS", "fp32 long-step thresholds met"


def compare(ref, got, area, klas, side, precision="fp64", step=1, nac=None):
    """Full report of `got` against `ref` (dicts with H U V Z, optionally W
    and F0..F3 or FLUX0..3)."""
    def norm(s):
        s = dict(s)
        for i in range(4):
            if f"F{i}" not in s and f"FLUX{i}" in s:
                s[f"F{i}"] = s[f"FLUX{i}"]
        if "W" not in s:
            s["W"] = np.sqrt(np.asarray(s["U"]) ** 2 + np.asarray(s["V"]) ** 2)
        return s

    ref, got = norm(ref), norm(got)
    stats = {}
    for f in STATE + FLUX:
        if f in ref and f in got:
            stats[f] = field_stats(f, ref[f], got[f], klas, nac)
    nonfinite = {f: int(np.count_nonzero(~np.isfinite(got[f]))) for f in STATE
                 if np.count_nonzero(~np.isfinite(got[f]))}
    cons = conservation_diffs(conservation(ref, area, klas, side), conservation(got, area, klas, side))
    v, why = verdict(precision, step, stats, cons, nonfinite)
    return {"precision": precision, "step": step, "verdict": v, "reason": why, "fields": stats,
            "conservation": cons}


def main(argv=None):
    import argparse

    from .mesh import make_config

    ap = argparse.ArgumentParser(description="compare two dumps of one seeded workload")
    ap.add_argument("ref")
    ap.add_argument("got")
    ap.add_argument("--config", required=True)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--step", type=int, default=1)
    a = ap.parse_args(argv)
    dt = np.float64 if a.precision == "fp64" else np.float32
    host = make_config(a.config, dtype=dt)
    rep = compare(read_dump(a.ref, dt), read_dump(a.got, dt), host.AREA, host.KLAS, host.SIDE, a.precision, a.step,
                  host.NAC)
    print(json.dumps(rep, default=float))
    return 0 if rep["verdict"] == "PASS" else 1


if __name__ == "__main__":
    sys.exit(main())
