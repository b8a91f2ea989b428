"""Throughput of the hydro-cal timestep on one MI355X.

Workload: the 207,234-cell floodplain mesh (20 km x 8 km, meandering channel,
wet/dry fronts), DT = 0.5 s, fp64 (headline) and fp32, synthetic seeded data.
A "step" is one full timestep (flux of every slot + update of every cell).

Prints ONE JSON line:
  value       cell-steps/s of the timed fp64 steps (state resident in HBM)
  roofline    the dominant kernel's algorithmic bytes / its launch duration
              (HIP events on the launch stream), against 8 TB/s
  cpu_baseline the reference's own CPU solver (hydro-cal-fortan/calculate.cpp,
              built unmodified into oracle/_ref/hydro_cc) timed on the same
              mesh, 1 host thread; the oracle (the C++ restatement of the
              reference kernels) timed beside it

The timed steps run, by default, as one PERSISTENT launch of the resident
TILES kernel: 1024-thread workgroups, one tile of ~810 cells per CU, each
workgroup keeping its tile on its CU for the whole launch.  --mode fused
times one launch per step instead, --block / --tile other tilings.

Usage: python bench.py [--gpus 1] [--steps 900] [--warmup 20] [--mode persistent]

The other headline measurements run after it and are carried in the same
line as sub-objects (--no-extras leaves them out):
  batched      128 uneven fp32 meshes (2K-207K cells, 500 steps) advanced as
               one batch, against the same meshes run one after another with
               plain launches; per-member accuracy against fp64 single runs
  adjoint      forward + adjoint of a 500-step gauge misfit (24,020 cells,
               checkpoints every 25 steps) against the plain forward run
  calibration  16 Manning hypotheses x 20 gradient steps on the 207,234-cell
               mesh over 900 steps, batched, fp32 trajectory
--ensemble, --adjoint and --calibrate print one of them as its own line.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hydro-bench_amd"))
sys.path.insert(0, ROOT)

# Algorithmic bytes per cell-step at 4 slots per cell (reads: state H U V Z,
# statics ZBC AREA FNC, per slot neighbour id 4 B + KLAS 1 B + COSF SINF SIDE;
# writes: H U V Z W):  24 s + 20 bytes.
BYTES_PER_CELL_STEP = {"f64": 24 * 8 + 20, "f32": 24 * 4 + 20}
HBM_PEAK_GBS = 8000.0
WORKLOAD = "floodplain_207234"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=900)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--mode", default="persistent",
                    choices=["persistent", "graph", "fused", "two_kernel", "graph_two_kernel"])
    ap.add_argument("--block", type=int, default=1024, help="threads per workgroup of the reordered-layout kernels")
    ap.add_argument("--launch", default="plain", choices=["plain", "cooperative"],
                    help="launch of the resident kernel (plain: ~10 us sooner to start; nothing else runs on the GPU)")
    ap.add_argument("--tile", type=int, default=-1,
                    help="cells per tile; -1: balanced, as many tiles as the GPU holds workgroups at once "
                         "(one per CU at --block 1024)")
    ap.add_argument("--regroup", type=int, default=0,
                    help="hydro_tune.regroup_steps: steps per launch and between regroupings of the task plan "
                         "by the wet/dry state (0 = the library's 200, -1 = one launch, no regrouping)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the forward line: no batched / adjoint / calibration sub-objects")
    ap.add_argument("--ensemble", action="store_true", help="batched-ensemble speedup instead of the headline line")
    ap.add_argument("--adjoint", action="store_true", help="adjoint cost ratio instead of the headline line")
    ap.add_argument("--adjoint-workload", default="floodplain_24020")
    ap.add_argument("--calibrate", action="store_true",
                    help="Manning calibration: 16 hypotheses x 20 gradient steps on the 207,234-cell mesh")
    ap.add_argument("--hypotheses", type=int, default=16)
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--calib-precision", default="fp32", choices=["fp32", "fp64"],
                    help="trajectory precision of --calibrate (the adjoint and the gradients are fp64 either way)")
    ap.add_argument("--calib-sequential", action="store_true",
                    help="--calibrate one hypothesis after another instead of all in one scenario batch")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 512 MiB device-to-device 5 times (a kernel of known traffic, to scale the "
                         "FETCH_SIZE / WRITE_SIZE counters by)")
    return ap.parse_args()


def copy_calibration():
    import torch

    n = 64 * 1024 * 1024   # 64 Mi float64 = 512 MiB per tensor
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 8, "copy_calibration_bytes_written": n * 8, "copies": 5}))


def time_gpu(host, mode, steps, warmup, block, tile=0, launch="cooperative", regroup=0):
    import torch

    import hydro_amd
    from hydro_amd.device import DeviceMesh

    tune = hydro_amd.default_tune(block=block, tile_cells=tile, resident_launch=1 if launch == "plain" else 0,
                                  regroup_steps=regroup)
    dm = DeviceMesh(host, device=0, tune=tune)
    tiles = dm.plan_stats()[0]
    init = host.state()
    if warmup:
        dm.step(min(warmup, host.max_steps), mode)
    torch.cuda.synchronize()
    dm.load_state(init)
    dm.set_time(0, 1)
    dm.step(0, mode)   # re-gather outside the timed region
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    dm.step(steps, mode)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1)
    launches = {"persistent": 1, "graph": steps, "fused": steps, "two_kernel": 2 * steps,
                "graph_two_kernel": 2 * steps}[mode]
    if mode == "persistent":   # one launch per regroup_steps steps at most
        launches = max(1, dm.last_timing()[1])
    st = dm.fetch("H")
    dm.close()
    return ms, wall, launches, st, tiles


def cpu_reference(host, steps):
    """The reference CPU solver on the same mesh (its input format allows no
    boundary series and an integer DT: DT = 1 s, discharge and level edges
    as walls); seconds per step from a 2-step and a (2+steps)-step run."""
    from oracle import cc_input

    per_step, runs = cc_input.time_reference(host, 2, 2 + steps)
    return host.CELL / per_step, per_step, runs


def cpu_baseline(host, budget_s, threads=1):
    from oracle import hydro_oracle

    old = hydro_oracle.set_threads(threads)
    try:
        r = hydro_oracle.OracleRun(host)
        t0 = time.perf_counter()
        r.step(1)
        one = time.perf_counter() - t0
        n = max(1, min(host.max_steps - 1, int(budget_s / max(one, 1e-6))))
        t0 = time.perf_counter()
        r.step(n)
        dt = time.perf_counter() - t0
    finally:
        hydro_oracle.set_threads(old)
    return host.CELL * n / dt, n, dt


def host_cpu():
    """CPU model, logical CPUs of the machine and CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0))
    return {"cpu": model, "nproc": os.cpu_count(), "usable": usable}


def cpu_threads():
    """Threads for the all-core CPU figure: the usable CPUs, capped by
    OMP_NUM_THREADS when it is set."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def _timed(fn):
    """Device time (ms) of the work fn issues on the current stream."""
    import torch

    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1)


def ensemble_bench(a):
    import numpy as np

    import hydro_amd
    import hydro_amd.mesh as M
    from hydro_amd.device import DeviceMesh
    from hydro_amd.ensemble import Batch, autotune_batch

    steps = 500
    t0 = time.perf_counter()
    # both precisions are built before the GPU is touched (the builder forks)
    h32 = M.make_ensemble(dtype=np.float32, threads=16)
    h64 = M.make_ensemble(dtype=np.float64, threads=16)
    cells = [h.CELL for h in h32]
    _log(f"ensemble: {len(h32)} meshes, {sum(cells)} cells, built in {time.perf_counter() - t0:.1f} s")
    tune = hydro_amd.default_tune(block=a.block)
    dms = [DeviceMesh(h, device=0, tune=tune) for h in h32]
    inits = [h.state() for h in h32]

    def reset(mode):
        for dm, st in zip(dms, inits):
            dm.load_state(st)
            dm.set_time(0, 1)
            if mode != "two_kernel":
                dm.step(0, mode)      # gather into the reordered layout outside the timed region
                dm.regroup()          # the plan grouped for this state, as a fresh prepare leaves it

    def sequential(mode):
        for dm in dms:
            dm.step(steps, mode)

    times = {}
    results = {}
    for mode in ("two_kernel", "fused"):
        for rep in range(2):          # the first run warms up
            reset(mode)
            times[f"sequential_{mode}"] = _timed(lambda: sequential(mode))
        for dm in dms:
            dm.sync()
        results[mode] = [dm.fetch("H", "U", "V", "Z") for dm in dms]
        _log(f"sequential {mode}: {times[f'sequential_{mode}']:.1f} ms")
    # the batch's workgroup size and mode: chosen by timing 20 steps of each
    (block, auto_mode), table = autotune_batch(dms, steps=20)
    _log(f"autotune: block {block}, {auto_mode} ({', '.join(f'{k[0]}/{k[1]} {v:.0f} us' for k, v in table.items())})")
    b = Batch(dms, tune=hydro_amd.default_tune(block=block))
    same = {}
    for mode in ("fused", "graph", "persistent"):
        for rep in range(2):
            reset("fused")
            b.step(0, mode)
            times[f"batched_{mode}"] = _timed(lambda: b.step(steps, mode))
        b.sync()
        got = [dm.fetch("H", "U", "V", "Z") for dm in dms]
        same[mode] = sum(all(np.array_equal(g[k], r[k]) for k in "HUVZ") for g, r in zip(got, results["two_kernel"]))
        _log(f"batched {mode}: {times[f'batched_{mode}']:.1f} ms, {same[mode]} members bit-equal to sequential")
    b.close()
    # accuracy of every member against its own fp64 single-mesh run
    worst_abs, worst_rel = 0.0, 0.0
    for i, h in enumerate(h64):
        ref = DeviceMesh(h, device=0)
        ref.step(steps, "fused")
        ref.sync()
        H64 = ref.fetch("H")["H"]
        ref.close()
        H32 = results["two_kernel"][i]["H"].astype(np.float64)
        d = np.abs(H32 - H64)
        worst_abs = max(worst_abs, float(d.max()))
        worst_rel = max(worst_rel, float(d.max() / max(np.abs(H64).max(), 1e-30)))
    for dm in dms:
        dm.close()
    best = min(("fused", "graph", "persistent"), key=lambda m: times[f"batched_{m}"])
    total = sum(cells) * steps
    out = {
        "metric": "batched speedup",
        "value": times["sequential_two_kernel"] / times[f"batched_{best}"],
        "unit": "x (sequential plain two-kernel launches / batched)",
        "n_gpus": 1, "steps": steps, "higher_is_better": True, "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (128 seeded floodplain meshes, seed 3003, cells log-uniform in [2000, 207234])",
        "config": {"workload": "ensemble_128", "members": len(cells), "cells": sum(cells), "min_cells": min(cells),
                   "max_cells": max(cells), "block": block, "batched_mode": best, "autotuned_mode": auto_mode},
        "autotune_us_per_step": {f"{k[0]}/{k[1]}": v for k, v in table.items()},
        "speedup_vs_sequential_fused": times["sequential_fused"] / times[f"batched_{best}"],
        "ms": times,
        "cell_steps_per_s": {k: total / (v / 1e3) for k, v in times.items()},
        "bit_equal_members": same,
        "fp32_vs_fp64_max_abs_dH_m": worst_abs,
        "fp32_vs_fp64_max_rel_dH": worst_rel,
    }
    return out


def adjoint_bench(a):
    import math

    import numpy as np
    import torch

    import hydro_amd
    from hydro_amd.adjoint import MisfitProblem, run_backward, run_forward
    from hydro_amd.device import DeviceMesh

    host = hydro_amd.make_config(a.adjoint_workload, dtype=np.float64)
    steps = min(500, host.max_steps)
    seg = 25 if steps == 500 else int(math.ceil(math.sqrt(steps)))
    rng = np.random.default_rng(4004)
    dm = DeviceMesh(host, device=0)
    dev = dm.device
    t = lambda x, dt=torch.float64: torch.tensor(x, dtype=dt, device=dev)
    level = host.Z.copy()
    gauges = rng.choice(np.nonzero(host.H > 0.5)[0], size=32, replace=False).astype(np.int32)
    prob = MisfitProblem(dm=dm, level0=t(level), U0=t(host.U), V0=t(host.V), gauges=t(gauges, torch.int32),
                         steps=steps, segment=seg)
    n = t(np.sqrt(host.FNC / 9.81))
    zbc = t(host.ZBC.copy())
    prob.z_obs = torch.zeros((steps, 32), dtype=torch.float64, device=dev)
    zg, _ = run_forward(prob, n * 1.05, zbc, keep_checkpoints=False)
    prob.z_obs = zg.clone()
    init = host.state()

    res = {}
    for rep in range(3):
        for mode in ("graph", "fused"):
            dm.load_state(init)
            dm.set_time(0, 1)
            dm.step(0, mode)
            res["forward" if mode == "fused" else "forward_graph"] = _timed(lambda: dm.step(steps, mode))
        box = {}
        res["misfit_forward"] = _timed(lambda: box.update(fb=run_forward(prob, n, zbc)))
        zg, buf = box["fb"]
        res["adjoint_sweep"] = _timed(lambda: box.update(g=run_backward(prob, buf, zg, n, zbc)))
    ratio = (res["misfit_forward"] + res["adjoint_sweep"]) / min(res["forward"], res["forward_graph"])
    g_n, g_z = box["g"]
    out = {
        "metric": "adjoint/forward cost ratio",
        "value": ratio,
        "unit": "x ((forward with checkpoints + recompute + adjoint) / plain forward)",
        "n_gpus": 1, "steps": steps, "higher_is_better": False, "vs_baseline": None, "dtype": "f64",
        "data": f"synthetic (seeded {host.CELL}-cell floodplain, 32 gauges, truth = Manning x 1.05)",
        "config": {"workload": a.adjoint_workload, "cells": host.CELL, "steps": steps, "checkpoint_every": seg},
        "ms": res,
        "ms_per_step": {k: v / steps for k, v in res.items()},
        "grad_norms": {"n": float(g_n.norm()), "zbc": float(g_z.norm())},
    }
    dm.close()
    return out


def calibrate_bench(a):
    """Manning calibration of the 207K floodplain: `hypotheses` starting
    fields, `iters` gradient steps each, all hypotheses in one scenario batch
    (default) or one after another, on an fp32 (default) or fp64
    trajectory; observations from the fp64 run of a truth field; every
    misfit reported is the fp64 forward run's (equal to the CPU oracle's)."""
    import numpy as np
    import torch

    import hydro_amd
    from hydro_amd.adjoint import MisfitProblem, run_forward
    from hydro_amd.calibrate import calibrate, calibrate_many, smooth_field
    from hydro_amd.device import DeviceMesh

    host = hydro_amd.make_config(WORKLOAD, dtype=np.float64)
    steps = min(900, host.max_steps)
    dm64 = DeviceMesh(host, device=0)
    dev = dm64.device
    t = lambda x, dt=torch.float64: torch.tensor(x, dtype=dt, device=dev)
    rng = np.random.default_rng(5005)
    level = host.Z.copy()
    level[host.H <= host.HM1] = host.ZBC[host.H <= host.HM1]
    gauges = rng.choice(np.nonzero(host.H > 0.5)[0], size=64, replace=False).astype(np.int32)

    def problem(dm):
        return MisfitProblem(dm=dm, level0=t(level), U0=t(host.U), V0=t(host.V), gauges=t(gauges, torch.int32),
                             steps=steps)

    p64 = problem(dm64)
    zbc = t(host.ZBC.copy())
    field = lambda seed: np.clip(smooth_field(host.xc, host.yc, seed, 0.04, 0.01, 2000.0), 0.005, 0.2)
    p64.z_obs = torch.zeros((steps, len(gauges)), dtype=torch.float64, device=dev)
    zg, _ = run_forward(p64, t(field(5005)), zbc, keep_checkpoints=False)
    p64.z_obs = zg.clone()
    if a.calib_precision == "fp32":
        dm = DeviceMesh(hydro_amd.make_config(WORKLOAD, dtype=np.float32), device=0)
        prob = problem(dm)
        prob.z_obs = p64.z_obs
    else:
        dm, prob = dm64, p64

    def misfit64(n):
        zg, _ = run_forward(p64, t(n), zbc, keep_checkpoints=False)
        return float(((zg - p64.z_obs) ** 2).sum())

    truth = field(5005)
    res = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if a.calib_sequential:
        cals = [calibrate(prob, t(field(6000 + h)), zbc, iters=a.iters) for h in range(a.hypotheses)]
    else:
        n0s = torch.stack([t(field(6000 + h)) for h in range(a.hypotheses)]).contiguous()
        cals = calibrate_many(prob, n0s, zbc, iters=a.iters)
    torch.cuda.synchronize()
    cal_wall = time.perf_counter() - t0
    for h, cal in enumerate(cals):
        n_final = cal.n.cpu().numpy()
        first64, final64 = misfit64(field(6000 + h)), misfit64(n_final)
        res.append({"hypothesis": h, "misfit_first": first64, "misfit_final": final64,
                    "trajectory_misfit_first": cal.losses[0], "trajectory_misfit_final": cal.best,
                    "rms_n_error_first": float(np.sqrt(np.mean((field(6000 + h) - truth) ** 2))),
                    "rms_n_error_final": float(np.sqrt(np.mean((n_final - truth) ** 2)))})
        _log(f"hypothesis {h}: fp64 misfit {first64:.3e} -> {final64:.3e}")
    ratios = [r["misfit_final"] / r["misfit_first"] for r in res]
    grads = a.hypotheses * a.iters
    out = {
        "metric": "calibration misfit reduction (final / first, fp64 forward)",
        "value": float(np.median(ratios)),
        "unit": "ratio (median over hypotheses)",
        "n_gpus": 1, "steps": steps, "higher_is_better": False, "vs_baseline": None,
        "dtype": "f32 trajectory, f64 adjoint and gradients" if a.calib_precision == "fp32" else "f64",
        "data": "synthetic (seeded 207,234-cell floodplain; truth and hypotheses smooth Manning fields, "
                "mean 0.04, std 0.01; 64 gauges)",
        "config": {"workload": WORKLOAD, "cells": host.CELL, "steps": steps, "checkpoint_every": prob.seg,
                   "hypotheses": a.hypotheses, "iterations": a.iters, "trajectory": a.calib_precision,
                   "batched": not a.calib_sequential},
        "wall_s": cal_wall,
        "ms_per_gradient": cal_wall * 1e3 / grads,
        "hypotheses": res,
    }
    if dm is not dm64:
        dm.close()
    dm64.close()
    return out


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the timestep is not sharded "
              f"in this build", file=sys.stderr)
        sys.exit(2)
    if a.ensemble:
        print(json.dumps(ensemble_bench(a)))
        return
    if a.adjoint:
        print(json.dumps(adjoint_bench(a)))
        return
    if a.calibrate:
        print(json.dumps(calibrate_bench(a)))
        return
    import numpy as np

    import hydro_amd

    host64 = hydro_amd.make_config(WORKLOAD, dtype=np.float64)
    steps = min(a.steps, host64.max_steps)
    ms, wall, launches, st, tiles = time_gpu(host64, a.mode, steps, a.warmup, a.block, a.tile, a.launch, a.regroup)
    if not np.all(np.isfinite(st["H"])):
        raise SystemExit("non-finite state after the timed steps")
    C = host64.CELL
    value = C * steps / (ms / 1e3)
    per_launch_ms = ms / launches
    two = a.mode in ("two_kernel", "graph_two_kernel")
    steps_per_launch = steps / launches if not two else 1.0
    bytes_per_launch = BYTES_PER_CELL_STEP["f64"] * C * (steps_per_launch if a.mode == "persistent" else 1)
    if two:
        bytes_per_launch = BYTES_PER_CELL_STEP["f64"] * C / 2.0
    achieved = bytes_per_launch / (per_launch_ms / 1e3) / 1e9
    # PERSISTENT runs the resident kernel whenever every tile fits on the GPU
    # at once (always at --block 1024 with balanced tiles), else the loop kernel
    resident = a.mode == "persistent" and tiles <= 256 * max(1, 1024 // a.block)
    kernel = {"persistent": "k_resident_tiles<double,0,false>" if resident else "k_persistent_tiles<double,0,false>",
              "graph": "k_fused_step_tick<double,0,false,0>",
              "fused": "k_fused_step<double,0,false,0>", "two_kernel": "k_calculate_flux+k_update_cell",
              "graph_two_kernel": "k_calculate_flux_tick+k_update_cell"}[a.mode]
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"{kernel.split('<')[0]}_{WORKLOAD}_f64_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        # a multi-step launch: its fixed part plus the per-step part for
        # this launch's steps (counter fit over two launch lengths)
        traffic = (tj["fixed_bytes"] + tj["bytes_per_step"] * steps_per_launch if "bytes_per_step" in tj
                   else tj.get("bytes_per_launch"))
    out = {
        "metric": "cell-steps/sec",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": ms / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 207,234-cell floodplain mesh, DT=0.5 s)",
        "config": {"workload": WORKLOAD, "cells": C, "edges": host64.n_edges, "dt": host64.DT,
                   "mode": a.mode, "block": a.block, "tiles": tiles,
                   "tile_cells": "balanced" if a.tile == -1 else (a.tile or a.block), "variant": "A",
                   "launch": a.launch if resident else None,
                   "regroup_steps": (200 if a.regroup == 0 else a.regroup) if a.mode != "two_kernel" else None},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_cell_step": BYTES_PER_CELL_STEP["f64"], "launches": launches,
                     "avg_launch_ms": per_launch_ms, "steps_per_launch": steps_per_launch},
        "wall_s": wall,
    }
    if not a.no_fp32:
        host32 = hydro_amd.make_config(WORKLOAD, dtype=np.float32)
        ms32, _, l32, _, _ = time_gpu(host32, a.mode, steps, a.warmup, a.block, a.tile, a.launch, a.regroup)
        v32 = C * steps / (ms32 / 1e3)
        out["fp32"] = {"value": v32, "ms_per_step": ms32 / steps,
                       "roofline_frac": v32 * BYTES_PER_CELL_STEP["f32"] / 1e9 / HBM_PEAK_GBS}
    if not a.no_cpu:
        v, n, dt = cpu_baseline(host64, a.cpu_seconds)
        port = {"value": v, "unit": "cell-steps/s", "cores": 1, "kind": "port",
                "sample": f"oracle (C++ restatement of the reference kernels, g++ -O2), "
                          f"{n} fp64 steps of the same 207,234-cell mesh in {dt:.1f} s, 1 thread"}
        nt = cpu_threads()
        if nt > 1:
            vo, no, dto = cpu_baseline(host64, a.cpu_seconds / 2, threads=nt)
            port["all_cores"] = {"value": vo, "unit": "cell-steps/s", "cores": nt, "kind": "port",
                                 "sample": f"the same oracle, slot and cell loops over {nt} OpenMP threads, "
                                           f"{no} fp64 steps in {dto:.1f} s"}
        port["host"] = host_cpu()
        out["cpu_baseline"] = port
        try:
            vr, per_step, runs = cpu_reference(host64, 60)
            out["cpu_baseline"] = {
                "value": vr, "unit": "cell-steps/s", "cores": 1, "kind": "reference",
                "sample": f"hydro-cal-fortan/calculate.cpp unmodified (g++ -O3), 60 fp64 steps of the same "
                          f"207,234-cell mesh, bed, Manning field and initial state at DT = 1 s with discharge "
                          f"and level edges as walls (its input format has no DT = 0.5 s and no safe boundary "
                          f"series); {per_step * 1e3:.1f} ms/step from runs of 2 and 62 steps "
                          f"({runs[2]:.1f} s, {runs[62]:.1f} s), 1 thread",
                "host": port["host"], "port": port}
        except (OSError, subprocess.SubprocessError) as e:
            print(f"reference CPU solver not timed ({e}); cpu_baseline is the oracle", file=sys.stderr)
    if not a.no_extras:
        out.update(extras(a))
    print(json.dumps(out))


def extras(a):
    """The other headline metrics, each as a compact sub-object of the line."""
    t0 = time.perf_counter()
    e = ensemble_bench(a)
    _log(f"batched: {time.perf_counter() - t0:.0f} s")
    best = e["config"]["batched_mode"]
    batched = {"metric": e["metric"], "value": e["value"], "unit": e["unit"], "higher_is_better": True,
               "dtype": e["dtype"], "config": e["config"],
               "speedup_vs_sequential_fused": e["speedup_vs_sequential_fused"],
               "ms": e["ms"], "batched_cell_steps_per_s": e["cell_steps_per_s"][f"batched_{best}"],
               "bit_equal_members": e["bit_equal_members"],
               "fp32_vs_fp64_max_abs_dH_m": e["fp32_vs_fp64_max_abs_dH_m"]}
    t0 = time.perf_counter()
    d = adjoint_bench(a)
    _log(f"adjoint: {time.perf_counter() - t0:.0f} s")
    adjoint = {"metric": d["metric"], "value": d["value"], "unit": d["unit"], "higher_is_better": False,
               "dtype": d["dtype"], "config": d["config"], "ms_per_step": d["ms_per_step"]}
    t0 = time.perf_counter()
    c = calibrate_bench(a)
    _log(f"calibration: {time.perf_counter() - t0:.0f} s")
    ratios = [h["misfit_final"] / h["misfit_first"] for h in c["hypotheses"]]
    calibration = {"metric": c["metric"], "value": c["value"], "unit": c["unit"], "higher_is_better": False,
                   "dtype": c["dtype"], "config": c["config"], "wall_s": c["wall_s"],
                   "ms_per_gradient": c["ms_per_gradient"], "max_ratio": max(ratios)}
    return {"batched": batched, "adjoint": adjoint, "calibration": calibration}


if __name__ == "__main__":
    main()
