#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the per-step SPH hot path on one MI355X.

Workload: 3D dam break, box 3.2 x 1.0 x 1.6 m, water 1.2 x 1.0 x 0.6 m,
dx = 0.0089 -> 1,013,040 fluid + 645,804 wall particles, WendlandC4 with
h = 1.3 dx, linear EOS, Shepard every 30 steps, FP32 pair arithmetic with
FP64 integration (FP32_BULK).  A "step" is one eachStep! of the whole scene
with the particle state resident in HBM.  The same scene in the plain FP64
form is timed beside it, and the FP64 CPU oracle is timed on a bounded sample
(a few steps of the same scene) on the host cores.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "ethersphcpu.jl_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

# Algorithmic bytes per fluid / wall particle of one launch of the force pass
# (forces + integration), 3D: FP64 form 128 B / 56 B, FP32_BULK 160 B / 28 B.
FORCE_BYTES = {0: (128.0, 56.0), 1: (160.0, 28.0)}
# Whole-step algorithmic bytes per fluid / wall particle-step, 3D, Shepard every 30.
STEP_BYTES = {0: (241.6, 56.0), 1: (232.8, 28.0)}
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_gpu(sc, prec, steps, warmup):
    from ethersph import _lib

    dev = sc.make_scene(prec)
    dev.step(warmup)
    dev.sync()
    dev.enable_timing(True)
    dev.phase_times()  # reset
    t0 = time.perf_counter()
    dev.step(steps)
    dev.sync()
    wall = time.perf_counter() - t0
    ph = dev.phase_times()
    ph["stats"] = dev.stats()
    n = dev.fluid_count()
    dev.close()
    return wall, ph, n


def cpu_baseline(sc, steps, threads):
    from oracle import oracle

    ref = oracle.from_scenario(sc)
    ref.step(1, threads)  # first step pays allocation
    t0 = time.perf_counter()
    ref.step(steps, threads)
    dt = time.perf_counter() - t0
    return sc.n_fluid * steps / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--dx", type=float, default=0.0089)
    ap.add_argument("--fp64-steps", type=int, default=300)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from ethersph import _lib, scenarios

    sc = scenarios.dam_break_3d(dx=a.dx, total_step=a.steps)
    nf, nw = sc.n_fluid, sc.n_wall
    log(f"workload dam_break_3d: {nf} fluid, {nw} wall; FP32_BULK {a.steps} steps")
    wall32, ph32, n_end = time_gpu(sc, _lib.FP32_BULK, a.steps, a.warmup)
    log(f"FP32_BULK: {wall32 * 1e3 / a.steps:.3f} ms/step, phases {ph32}")
    wall64, ph64, _ = time_gpu(sc, _lib.FP64, a.fp64_steps, a.warmup)
    log(f"FP64: {wall64 * 1e3 / a.fp64_steps:.3f} ms/step, phases {ph64}")

    value = nf * a.steps / wall32
    value64 = nf * a.fp64_steps / wall64
    fb, wb = FORCE_BYTES[1]
    t_force = ph32["forces"] * 1e-3
    force_bytes = fb * nf + wb * nw
    achieved = force_bytes / t_force / 1e9
    sb, swb = STEP_BYTES[1]
    step_gbs = (sb * nf + swb * nw) * a.steps / wall32 / 1e9

    cpu = None
    if not a.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"CPU baseline: FP64 oracle, {a.cpu_steps} steps of the full scene on {threads} threads")
        v, dt = cpu_baseline(sc, a.cpu_steps, threads)
        cpu = {"value": v, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": f"{a.cpu_steps} steps of the full {nf}-particle scene, FP64 oracle (OpenMP), {dt:.1f} s"}

    out = {
        "metric": "particle-steps/s at 1 GPU + %HBM peak; ensemble scenes/s, p50/p99 step latency",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall32 * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair arithmetic, f64 integration",
        "data": "synthetic (3D dam-break lattice, seeded generator)",
        "config": {"workload": "3D dam break, 1,013,040 fluid + 645,804 wall particles, WendlandC4 h=1.3dx, "
                               "FP32_BULK, Shepard every 30 steps",
                   "n_fluid": nf, "n_wall": nw, "n_fluid_end": n_end, "dx": a.dx},
        "roofline": {"bound": "hbm", "kernel": "k_forces", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "step_algorithmic_GBps": step_gbs,
        "phases_ms": {"fp32_bulk": ph32, "fp64": ph64},
        "fp64_form": {"value": value64, "ms_per_step": wall64 * 1e3 / a.fp64_steps, "steps": a.fp64_steps},
        "fp32_over_fp64": value / value64,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
