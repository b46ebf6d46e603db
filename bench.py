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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cells", type=int, default=4_000_000)
    ap.add_argument("--particles", type=int, default=500_000)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bounded CPU baseline sample length")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true")
    return ap.parse_args()


def cpu_baseline(world, seconds):
    from oracle.pyoracle import OracleWorld

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)), 16))
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
        "kind": "port",
        "sample": f"oracle FP32 OpenMP, same {n}-cell world, steps 2..{steps + 1} after 1 warm-up step "
                  f"({el:.1f} s)",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (--gpus {args.gpus} requested); "
              "independent worlds would shard one world per GPU with no exchange.", file=sys.stderr)
        sys.exit(2)

    from alien_amd import engine, worlds

    world = worlds.large_fluid(size=(args.size, args.size), num_cells=args.cells, num_particles=args.particles,
                               steps=args.steps)
    n = len(world.cells)
    sim = engine.Simulation.from_world(world)
    device = engine.gpu_name()
    sim.step(args.warmup)

    t0 = time.perf_counter()
    sim.step(args.steps)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    value = n * args.steps / elapsed
    live_particles = sim.counts()[1]

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
                    traffic = json.load(f).get("bytes_per_launch")
            roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                        "kernel": "k_neighbor", "avg_launch_us": avg_s * 1e6,
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
                   "device": device},
        "roofline": roofline,
        "step_bandwidth_GBs": step_bytes / (elapsed / args.steps) / 1e9,
        "live_particles_after": live_particles,
        "kernels": kernels,
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(world, args.cpu_seconds)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
