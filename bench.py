#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the per-step SPH hot path on one MI355X.

Workload: 3D dam break, box 3.2 x 1.0 x 1.6 m, water 1.2 x 1.0 x 0.6 m,
dx = 0.0089 -> 1,013,040 fluid + 645,804 wall particles, WendlandC4 with
h = 1.3 dx, linear EOS, Shepard every 30 steps, FP32 pair arithmetic with
FP64 integration (FP32_BULK).  A "step" is one eachStep! of the whole scene
with the particle state resident in HBM.  The same scene in the plain FP64
form is timed beside it over the same window (fp32_over_fp64), and both forms
again over the workload's 1,000 steps when --steps is another count
(long_run); the FP64 CPU oracle is timed on a bounded sample (a few steps of
the same scene) on every host core available to the process.

Two further legs report the other halves of the metric: "ensemble" (256 2D
scenes, half lid-driven cavities and half Poiseuille channels, sizes
log-uniform in [2k, 60k], batched with graphs versus launched one scene at a
time) and "serving" (Poisson arrivals, faster than the server completes them,
of 2D dam breaks, 3D Poiseuille pipes and 2D thermal cavities at the
workload's sizes and step counts under a 200 GB HBM budget: completed
scenes/s and p50/p99 step latency).  Both are bounded samples, stated in
their output.

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
# (forces + integration), 3D, by the survey's counting rule (every
# per-particle array the phase must touch, read once and written once;
# neighbour re-reads are cache traffic): FP64 form 128 B / 56 B, FP32_BULK
# 160 B / 28 B.
FORCE_BYTES = {0: (128.0, 56.0), 1: (160.0, 28.0)}
# Whole-step algorithmic bytes per fluid / wall particle-step, 3D, Shepard every 30.
STEP_BYTES = {0: (241.6, 56.0), 1: (232.8, 28.0)}
HBM_PEAK_GBS = 8000.0


def measured_traffic(kernel, prec="fp32_bulk"):
    """HBM bytes per dispatch from the committed rocprofv3 counter passes
    (calibrated FETCH_SIZE + WRITE_SIZE), or None when not profiled."""
    try:
        with open(os.path.join(ROOT, "profiles", "dam_break_3d_traffic.json")) as fh:
            return float(json.load(fh)[kernel][prec])
    except (OSError, KeyError, ValueError):
        return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_gpu(sc, prec, steps, warmup, phase_steps=200):
    """Steps of one scene replayed from its captured step graph (the timed
    region: `steps` steps with the state resident in HBM), then `phase_steps`
    more with plain launches and per-phase HIP events for the breakdown."""
    from ethersph import _lib

    dev = sc.make_scene(prec)
    dev.enable_graph(True)
    dev.step(warmup)
    dev.sync()
    t0 = time.perf_counter()
    dev.step(steps)
    dev.sync()
    wall = time.perf_counter() - t0
    dev.enable_graph(False)
    dev.enable_timing(True)
    dev.phase_times()  # reset
    dev.step(phase_steps)
    dev.sync()
    ph = dev.phase_times()
    ph["stats"] = dev.stats()
    tuned = dev.ctx.tuned_depth(sc.dim, sc.kernel.kind, prec, 1)
    if tuned is not None:
        ph["gather_depth"] = {"chosen": tuned[0], "forces_ms_at_depth_1_2_4": tuned[1]}
    wg = dev.ctx.tuned_workgroup(sc.dim, sc.kernel.kind, prec, 1)
    if wg is not None:
        ph["workgroup"] = {"chosen": wg[0], "forces_ms_at_64_128_256_threads": wg[1]}
    # each pass's own choice (continuity timed dry, Shepard whether due or not)
    for q, name in ((1, "continuity"), (2, "shepard")):
        t = dev.ctx.tuned_pass(sc.dim, sc.kernel.kind, prec, q, 1)
        if t is not None:
            ph[f"{name}_tuning"] = {"depth": t[0], "threads": t[1],
                                    "ms_at_depth_1_2_4_x_threads_64_128_256": [round(v, 4) for v in t[2]]}
    n = dev.fluid_count()
    dev.close()
    return wall, ph, n


def ensemble_leg(steps, alone_steps, n_scenes, full_steps=5000):
    """Batched ensemble of 2D scenes (half lid-driven cavities, half
    Poiseuille channels) against the same scenes launched one at a time.
    scenes/s is quoted for scenes of full_steps steps from the measured
    per-step times."""
    from ethersph import _lib, scenarios
    from ethersph.solve import Batch

    scs = scenarios.ensemble_2d(n_scenes=n_scenes, seed=0, total_step=full_steps)
    scenes = [sc.make_scene(_lib.FP32_BULK) for sc in scs]
    npart = sum(sc.n_fluid for sc in scs)
    b = Batch(scenes, graph=True)
    b.step(5)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    t_batch = (time.perf_counter() - t0) / steps
    b.close()
    # one at a time: each scene alone, in turn -- with plain launches, then
    # with the scene's own captured step graph
    t_plain = t_alone = float("nan")
    if alone_steps > 0:
        for s in scenes:
            s.step(1)
        scenes[0].sync()
        t0 = time.perf_counter()
        for s in scenes:
            s.step(alone_steps)
            s.sync()
        t_plain = (time.perf_counter() - t0) / alone_steps
        # with the scene's own captured step graphs, one scene at a time: each
        # scene captures its graphs in an untimed run as long as the timed one
        # (single steps, with and without the Shepard pass, and runs of 2-16
        # steps), runs its timed steps and releases its graphs before the next
        # (at most six live graph executables)
        t_alone = 0.0
        for s in scenes:
            s.enable_graph(True)
            s.step(alone_steps)
            s.sync()
            t0 = time.perf_counter()
            s.step(alone_steps)
            s.sync()
            t_alone += time.perf_counter() - t0
            s.enable_graph(False)
        t_alone /= alone_steps
    for s in scenes:
        s.close()
    batched = n_scenes / (t_batch * full_steps)
    alone = n_scenes / (t_alone * full_steps)
    plain = n_scenes / (t_plain * full_steps)
    return {"scenes": n_scenes, "fluid_particles": npart, "steps_per_scene": full_steps,
            "batched_ms_per_step": t_batch * 1e3, "one_at_a_time_ms_per_step": t_plain * 1e3,
            "one_at_a_time_graph_ms_per_step": t_alone * 1e3,
            "batched_scenes_per_s": batched, "one_at_a_time_scenes_per_s": plain,
            "one_at_a_time_graph_scenes_per_s": alone, "speedup": batched / plain,
            "speedup_vs_graphed_scenes": batched / alone,
            "batched_particle_steps_per_s": npart / t_batch,
            "sample": f"{steps} batched steps (step graphs, up to 16 steps in one) and {alone_steps} steps per scene alone "
                      f"(plain launches, then the scene's own graph), FP32_BULK; scenes/s extrapolated to "
                      f"{full_steps}-step scenes"}


def serving_leg(n_requests, rate, step_scale, budget_gb, advance_steps):
    """Poisson arrivals of mixed requests served by continuous batching.  The
    arrival rate is set above the server's capacity (the queue never runs
    dry), so completed scenes/s measures capacity, not the arrival rate."""
    from ethersph import _lib, scenarios
    from ethersph.serving import Server

    lo, hi = max(1, int(1000 * step_scale)), max(1, int(10000 * step_scale))
    stream = scenarios.serving_stream(n_requests=n_requests, seed=1, rate_per_s=rate, steps=(lo, hi))
    ctx = _lib.Context(0, int(budget_gb * 1e9))
    sv = Server(ctx)
    ids = []
    t0 = time.perf_counter()
    k = 0
    psteps = 0
    while True:
        now = time.perf_counter() - t0
        while k < len(stream) and stream[k][0] <= now:
            _, sc, n = stream[k]
            ids.append(sv.submit_scenario(sc, _lib.FP32_BULK, n))
            psteps += sc.n_fluid * n
            k += 1
        st = sv.stats()
        if k == len(stream) and st["running"] == 0 and st["queued"] == 0:
            break
        if st["running"] == 0 and st["queued"] == 0:
            time.sleep(max(0.0, stream[k][0] - now))
            continue
        sv.advance(advance_steps)
    wall = time.perf_counter() - t0
    st = sv.stats()
    failed = sum(1 for r in ids if sv.status(r)[0] == _lib.REQ_FAILED)
    sv.close()
    ctx.close()
    return {"requests": n_requests, "completed": int(st["completed"]), "failed": failed,
            "scenes_per_s": st["completed"] / wall, "p50_step_ms": st["p50_step_ms"],
            "p99_step_ms": st["p99_step_ms"], "particle_steps_per_s": psteps / wall, "wall_s": wall,
            "last_arrival_s": stream[-1][0], "batch_steps": int(st["batch_steps"]),
            "sample": f"{n_requests} Poisson arrivals at {rate}/s (all queued within "
                      f"{stream[-1][0]:.2f} s, above capacity), steps {lo}-{hi} per request "
                      f"(the workload's 1,000-10,000 x {step_scale}), sizes as the workload names them, "
                      f"{budget_gb} GB budget, FP32_BULK"}


def host_cores():
    """Threads for the CPU baseline: every core this process may run on
    (sched affinity), or OMP_NUM_THREADS when the host sets it lower; with
    the host's core count and CPU model for the record."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(avail, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else avail
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, {"nproc": os.cpu_count(), "affinity_cores": avail, "OMP_NUM_THREADS": omp, "cpu_model": model}


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
    ap.add_argument("--long-steps", type=int, default=1000,
                    help="both forms over this many steps for the fp32_over_fp64 figure of the configured run")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble-scenes", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=300)
    ap.add_argument("--ensemble-alone-steps", type=int, default=30)
    ap.add_argument("--serving-requests", type=int, default=16)
    ap.add_argument("--serving-rate", type=float, default=50.0)
    ap.add_argument("--serving-step-scale", type=float, default=1.0)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-only", action="store_true", help="run only the ensemble leg (for profiling)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from ethersph import _lib, scenarios

    if a.ensemble_only:
        print(json.dumps({"ensemble": ensemble_leg(a.ensemble_steps, a.ensemble_alone_steps, a.ensemble_scenes)}))
        return

    sc = scenarios.dam_break_3d(dx=a.dx, total_step=a.steps)
    nf, nw = sc.n_fluid, sc.n_wall
    log(f"workload dam_break_3d: {nf} fluid, {nw} wall; FP32_BULK {a.steps} steps")
    wall32, ph32, n_end = time_gpu(sc, _lib.FP32_BULK, a.steps, a.warmup)
    log(f"FP32_BULK: {wall32 * 1e3 / a.steps:.3f} ms/step, phases {ph32}")
    # the FP64 form over the same window (warmup, then the same steps)
    wall64, ph64, _ = time_gpu(sc, _lib.FP64, a.steps, a.warmup)
    log(f"FP64: {wall64 * 1e3 / a.steps:.3f} ms/step, phases {ph64}")
    value = nf * a.steps / wall32
    value64 = nf * a.steps / wall64
    # both forms over the configured run's length (the workload's 1,000 steps)
    longrun = None
    if a.long_steps > 0 and a.long_steps != a.steps:
        l32, _, _ = time_gpu(sc, _lib.FP32_BULK, a.long_steps, a.warmup, phase_steps=0)
        l64, _, _ = time_gpu(sc, _lib.FP64, a.long_steps, a.warmup, phase_steps=0)
        longrun = {"steps": a.long_steps, "warmup": a.warmup, "fp32_bulk_ms_per_step": l32 * 1e3 / a.long_steps,
                   "fp64_ms_per_step": l64 * 1e3 / a.long_steps, "fp32_over_fp64": l64 / l32,
                   "fp32_bulk_particle_steps_per_s": nf * a.long_steps / l32}
        log(f"{a.long_steps}-step run: {longrun}")
    fb, wb = FORCE_BYTES[1]
    t_force = ph32["forces"] * 1e-3
    force_bytes = fb * nf + wb * nw
    achieved = force_bytes / t_force / 1e9
    sb, swb = STEP_BYTES[1]
    step_gbs = (sb * nf + swb * nw) * a.steps / wall32 / 1e9

    cpu = None
    if not a.no_cpu_baseline:
        threads, host = host_cores()
        log(f"CPU baseline: FP64 oracle, {a.cpu_steps} steps of the full scene on {threads} threads, {host}")
        v, dt = cpu_baseline(sc, a.cpu_steps, threads)
        cpu = {"value": v, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": f"{a.cpu_steps} steps of the full {nf}-particle scene (after one untimed step), "
                         f"FP64 oracle (C, OpenMP), {dt:.1f} s",
               **host}

    ens = None
    if not a.no_ensemble:
        log(f"ensemble: {a.ensemble_scenes} 2D scenes, batched vs one at a time")
        try:
            ens = ensemble_leg(a.ensemble_steps, a.ensemble_alone_steps, a.ensemble_scenes)
        except Exception as e:  # reported in the JSON line, the headline stands
            ens = {"error": repr(e)}
        log(f"ensemble: {ens}")
    srv = None
    if not a.no_serving:
        log(f"serving: {a.serving_requests} requests")
        try:
            srv = serving_leg(a.serving_requests, a.serving_rate, a.serving_step_scale, 200.0, 50)
        except Exception as e:
            srv = {"error": repr(e)}
        log(f"serving: {srv}")

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
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("k_forces"),
                     "algorithmic_bytes": force_bytes},
        "step_algorithmic_GBps": step_gbs,
        "phases_ms": {"fp32_bulk": ph32, "fp64": ph64},
        "fp64_form": {"value": value64, "ms_per_step": wall64 * 1e3 / a.steps, "steps": a.steps,
                      "warmup": a.warmup},
        "fp32_over_fp64": value / value64,
        "long_run": longrun,
        "cpu_baseline": cpu,
        "ensemble": ens,
        "serving": srv,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
