"""Benchmark of the physics timestep on one MI355X.

Workload (default): the large fluid world — toroidal 4096x4096, 4,000,000 unconnected cells with
velocity N(0, 0.1^2), 500,000 energy particles, SPH motion with smoothing length 0.65 (interaction
radius 1.3), FP32.  A "step" is one physics timestep of that world.  The metric is
cell-timesteps/s = cells x timed steps / wall time of the timed steps, with all state resident in
HBM (the step call returns after the device finished).

Output: one JSON line (bench contract) with two extra objects:
  roofline     the dominant kernel (the neighbour pass) against HBM bandwidth; its duration comes
               from HIP events recorded around every launch on the engine's stream during a
               second, identical measurement region of the same number of steps;
  cpu_baseline the CPU oracle (FP32, OpenMP) on the host cores over a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak
# algorithmic (compulsory) bytes per unit and step: fluid cell 56 B, energy particle 52 B
BYTES_PER_FLUID_CELL = 56
BYTES_PER_PARTICLE = 52
NEIGHBOR_KERNEL = "k_neighbor_cells"  # the neighbour pass of fluid worlds (the dominant kernel)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 1000 for the large fluid world, 500 per call for the batch)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cells", type=int, default=4_000_000)
    ap.add_argument("--particles", type=int, default=500_000)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bounded CPU baseline sample length")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--workload", choices=["large_fluid", "batch", "serving"], default="large_fluid",
                    help="large_fluid: the headline line; batch: 256 uneven worlds in one graph vs per-world "
                         "sequential runs; serving: Poisson job arrivals through the job server")
    ap.add_argument("--batch-worlds", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=500,
                    help="steps per call of the batched-worlds leg of the default line (config 3: 500)")
    ap.add_argument("--radiation", action="store_true",
                    help="large fluid world with ALIEN's default cell radiation (probability 0.03, age strength "
                         "0.00002 per colour): the cost of the process, not the headline workload")
    ap.add_argument("--no-batch", action="store_true",
                    help="default line without its batched-worlds leg (batched vs per-world sequential)")
    ap.add_argument("--serving-seconds", type=float, default=60.0, help="arrival window of the serving run")
    ap.add_argument("--serving-rate", type=float, default=20.0)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of known size (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 500 if args.workload == "batch" else 1000
    return args


def copy_calibration():
    """A plain copy kernel moving a known number of bytes, for scaling rocprofv3 FETCH_SIZE /
    WRITE_SIZE readings into HBM bytes (read 256 MiB, write 256 MiB per launch)."""
    import torch

    n = 64 * 1024 * 1024  # float32 elements = 256 MiB
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 5}}), flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(world, seconds):
    from oracle.pyoracle import OracleWorld

    # every core this process may run on, unless OMP_NUM_THREADS says fewer (a shared host's
    # share of cores)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
    cores = max(1, cores)
    o = OracleWorld(world.size, world.params, precision=32, threads=cores)
    o.set_data(world.cells, world.particles)
    o.step(1)
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step(1)
        steps += 1
        el = time.perf_counter() - t0
        if (el >= seconds and steps >= 2) or steps >= 1000:
            break
    n = len(world.cells)
    return {
        "value": n * steps / el,
        "unit": "cell-timesteps/s",
        "cores": cores,
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": f"oracle FP32 OpenMP, same {n}-cell world, steps 2..{steps + 1} after 1 warm-up step "
                  f"({el:.1f} s)",
    }


def progress(msg):
    print(msg, file=sys.stderr, flush=True)


def gen_processes():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def batch_vs_sequential(count, steps, warmup):
    """256 worlds of uneven size (side log-uniform in [256, 2048], density U(0.02, 0.3), soft-body
    patches, collision mode) advanced together, one captured graph replay per step, against the
    same worlds run one after another as plain single-world simulations of the same library."""
    from alien_amd import engine, worlds

    t0 = time.perf_counter()
    ws = worlds.batched_worlds(count=count, steps=steps, processes=gen_processes())
    gen_s = time.perf_counter() - t0
    cells = sum(len(w.cells) for w in ws)
    progress(f"generated {len(ws)} worlds, {cells} cells in {gen_s:.1f} s")
    b = engine.Batch()
    for w in ws:
        b.add(w)
    b.step(warmup)
    tune = b.autotune(trial_steps=max(2, min(10, steps)))
    t0 = time.perf_counter()
    b.step(steps)
    batched_s = time.perf_counter() - t0
    b.close()
    progress(f"batched: {steps} steps in {batched_s:.2f} s (autotune block {tune[0]}, strip cap {tune[1]}, "
             f"chunk rows {tune[2]})")
    # per-world sequential: each world alone, same steps, same library
    seq_s = 0.0
    for k, w in enumerate(ws):
        sim = engine.Simulation.from_world(w)
        sim.step(warmup)
        t0 = time.perf_counter()
        sim.step(steps)
        seq_s += time.perf_counter() - t0
        sim.close()
        if (k + 1) % 32 == 0:
            progress(f"sequential: {k + 1} worlds, {seq_s:.2f} s of stepping")
    value = cells * steps / batched_s
    seq_value = cells * steps / seq_s
    return {
        "value": value, "ms_per_step": 1000.0 * batched_s / steps, "steps": steps, "warmup": warmup,
        "worlds": len(ws), "cells": cells,
        "sides": [int(min(w.size[0] for w in ws)), int(max(w.size[0] for w in ws))],
        "seed": worlds.SEED_BASE + 3,
        "per_world_sequential": {"value": seq_value, "seconds": seq_s, "ms_per_step": 1000.0 * seq_s / steps},
        "batched_over_sequential": value / seq_value,
        "autotune": {"block": tune[0], "strip_cap": tune[1], "chunk_rows": tune[2], "trial_ms_per_step": tune[3]},
        "generation_s": gen_s,
    }


def bench_batch(args):
    from alien_amd import engine

    r = batch_vs_sequential(args.batch_worlds, args.steps, args.warmup)
    print(json.dumps({
        "metric": "cell-timesteps/s (256 uneven worlds, batched, 1 GPU)", "value": r["value"],
        "unit": "cell-timesteps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched_worlds", "worlds": r["worlds"], "cells": r["cells"], "sides": r["sides"],
                   "seed": r["seed"], "device": engine.gpu_name()},
        "per_world_sequential": r["per_world_sequential"],
        "batched_over_sequential": r["batched_over_sequential"],
        "autotune": r["autotune"],
        "generation_s": r["generation_s"],
    }), flush=True)


def bench_serving(args):
    """Poisson job arrivals (rate, window) of batched-workload worlds with log-uniform step
    counts, admitted by an HBM budget into the continuous-batching server; latency = submit to
    snapshot ready."""
    from alien_amd import abi, engine, worlds

    t0 = time.perf_counter()
    jobs = worlds.serving_jobs(rate=args.serving_rate, duration=args.serving_seconds, processes=gen_processes())
    gen_s = time.perf_counter() - t0
    progress(f"generated {len(jobs)} jobs in {gen_s:.1f} s")
    srv = engine.Server(budget_bytes=int(args.budget_gb * 2**30), chunk_steps=50)
    ids = []
    start = time.perf_counter()
    for t_arr, w in jobs:
        delay = t_arr - (time.perf_counter() - start)
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(w.size, w.params, w.cells, w.particles, w.steps))
        if len(ids) % 100 == 0:
            progress(f"submitted {len(ids)} jobs")
    pending, last, t_wait = set(ids), time.perf_counter(), time.perf_counter()
    while pending:
        time.sleep(1.0)
        for j in list(pending):
            st = srv.poll(j).state
            if st == abi.JOB_FAILED:
                raise RuntimeError(f"serving job {j} failed")
            if st == abi.JOB_DONE:
                pending.discard(j)
        now = time.perf_counter()
        if now - last >= 20.0:
            progress(f"{len(ids) - len(pending)} of {len(ids)} jobs done")
            last = now
        if now - t_wait > 3600.0:
            raise RuntimeError("serving jobs did not finish within an hour")
    infos = [srv.poll(j) for j in ids]
    lat = sorted(i.finishTime - i.submitTime for i in infos)
    work = sum(i.numCells * i.stepsDone for i in infos)
    span = max(i.finishTime for i in infos) - min(i.submitTime for i in infos)
    srv.close()

    def pct(q):
        return lat[min(len(lat) - 1, int(q * (len(lat) - 1) + 0.5))]

    print(json.dumps({
        "metric": "serving job latency p50/p99 and aggregate cell-timesteps/s (1 GPU)", "value": work / span,
        "unit": "cell-timesteps/s", "n_gpus": 1, "steps": None, "warmup": 0, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "serving", "rate_per_s": args.serving_rate, "window_s": args.serving_seconds,
                   "jobs": len(jobs), "budget_gb": args.budget_gb, "seed": worlds.SEED_BASE + 4,
                   "device": engine.gpu_name()},
        "latency_s": {"p50": pct(0.5), "p99": pct(0.99), "max": lat[-1]},
        "generation_s": gen_s,
    }), flush=True)


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (--gpus {args.gpus} requested); "
              "independent worlds would shard one world per GPU with no exchange.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "serving":
        bench_serving(args)
        return

    from alien_amd import engine, worlds

    world = worlds.large_fluid(size=(args.size, args.size), num_cells=args.cells, num_particles=args.particles,
                               steps=args.steps)
    if args.radiation:
        for c in range(7):
            world.params.radiationCellAgeStrength[c] = 0.00002
    n = len(world.cells)
    t0 = time.perf_counter()
    sim = engine.Simulation.from_world(world)  # host arrays -> device (set_data)
    set_s = time.perf_counter() - t0
    device = engine.gpu_name()
    sim.step(args.warmup)

    t0 = time.perf_counter()
    sim.step(args.steps)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    value = n * args.steps / elapsed
    live_particles = sim.counts()[1]
    t0 = time.perf_counter()
    sim.get_data()  # device -> host arrays
    get_s = time.perf_counter() - t0

    roofline = None
    kernels = {}
    if not args.no_kernel_timing:
        sim.set_kernel_timing(True)
        sim.step(args.steps)
        times = sim.kernel_times()
        sim.set_kernel_timing(False)
        kernels = {k: {"launches": v[0], "avg_us": 1000.0 * v[1] / max(v[0], 1)} for k, v in times.items()}
        nb = times.get("neighbor")
        if nb and nb[0]:
            avg_s = nb[1] / nb[0] / 1000.0
            per_launch = BYTES_PER_FLUID_CELL * n
            achieved = per_launch / avg_s / 1e9
            traffic = None
            tpath = os.path.join(ROOT, "profiles", "traffic_neighbor_large_fluid.json")
            if os.path.exists(tpath):
                with open(tpath) as f:
                    rec = json.load(f)
                # (counter readings of the same kernel only)
                if rec.get("kernel") == NEIGHBOR_KERNEL:
                    traffic = rec.get("bytes_per_launch")
            roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                        "kernel": NEIGHBOR_KERNEL, "avg_launch_us": avg_s * 1e6,
                        "bytes_per_launch": per_launch}

    step_bytes = BYTES_PER_FLUID_CELL * n + BYTES_PER_PARTICLE * args.particles
    result = {
        "metric": "cell-timesteps/s (4M cells, 1 GPU)",
        "value": value,
        "unit": "cell-timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "large_fluid", "world": f"{args.size}x{args.size}", "cells": n,
                   "particles": args.particles, "motion": "fluid (SPH, radius 1.3)", "seed": world.seed,
                   "device": device, "cell_radiation": "ALIEN defaults" if args.radiation else "off"},
        "roofline": roofline,
        # the whole step against HBM: algorithmic bytes of all cells and particles per step
        "step_bandwidth_GBs": step_bytes / (elapsed / args.steps) / 1e9,
        "step_frac": step_bytes / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS,
        # the host boundary: one set_data before and one get_data after the run (not per step)
        "host_boundary": {"set_data_s": set_s, "get_data_s": get_s,
                          "value_with_set_and_get": n * args.steps / (elapsed + set_s + get_s)},
        "live_particles_after": live_particles,
        "kernels": kernels,
    }
    sim.close()
    if not args.no_cpu_baseline:
        progress("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(world, args.cpu_seconds)
    else:
        result["cpu_baseline"] = None
    if not args.no_batch:
        # the 256-world batch (500 steps per call, its own seed) against per-world sequential runs
        progress("batched worlds")
        b = batch_vs_sequential(args.batch_worlds, args.batch_steps, 10)
        result["batch"] = b
        result["batched_over_sequential"] = b["batched_over_sequential"]
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
