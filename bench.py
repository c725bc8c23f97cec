"""Benchmark: agent-steps/s of 1M-agent boids2d on one MI355X, FP64 and FP32.

Workload (1,048,576 boids, the metric's configuration): examples/boids2d_cfg.abl
with agent_density 0.065536 (a 4000 x 4000 box), positions uniform, velocities
uniform in [-1,1]^2, interaction radius 10, separation radius 2, weights
separation/alignment/cohesion 1.5/1.0/1.0, max speed 2; initial state from the
model's own main() and the reference's fixed xorshift128+ state.  A "step" is
one timestep of simulate(): the grid rebuild (bin, scan, scatter, stable
rank+gather) plus the update_boid neighbour kernel, agents resident in HBM.

The JSON line carries
  value            agent-steps/s, FP64, whole timestep (grid + neighbour step);
  fp32             the same at use_float precision;
  roofline         the neighbour kernel: algorithmic bytes per launch
                   (N * 64 B + 4 * (cells + 1) for FP64) / its average duration,
                   timed with HIP events on the simulation's own stream;
  cpu_baseline     the reference C path (oracle/_ref: the restated generated
                   program linked against the reference's libabl.c, gcc -O2
                   -fopenmp, brute-force near loop) on a bounded sample: the
                   first M agents' update_boid against the full 1M population.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
N_AGENTS = 1 << 20
DENSITY = 0.065536
PARAMS = dict(num_agents=N_AGENTS, agent_density=DENSITY, interaction_radius=10.0,
              separation_radius=2.0, steer_scale=1.0, collision_scale=1.5, match_scale=1.0,
              max_speed=2.0, time_scale=0.0005, global_scale=0.15)


def neighbour_bytes(n: int, precision: str, cells: int) -> int:
    """Algorithmic HBM bytes of one update_boid launch: every agent's pos and
    velocity read once and written once, plus the cell offsets."""
    fb = 8 if precision == "f64" else 4
    return n * 4 * fb * 2 + 4 * (cells + 1)


def run_gpu(precision: str, steps: int, warmup: int, timing_steps: int):
    import openabl_amd
    from openabl_amd import Simulation, load_library
    lib = load_library()
    lib.abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    sim = Simulation("boids2d", precision, params=PARAMS)
    sim.init_agents()
    sim.simulate(warmup)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.simulate(steps)
    sim.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    # kernel durations: event pairs around every launch on the sim's stream
    sim.set_config("hip.timing", 1)
    sim.simulate(timing_steps)
    total_ms, nbr_ms, grid_ms = sim.last_timing()
    box = sim.get_param("max_pos")
    cells = int(-(-box // (10.0 * (1 + 1 / 1048576)))) ** 2
    sim.close()
    return {
        "wall_s": wall,
        "agent_steps_per_s": N_AGENTS * steps / wall,
        "ms_per_step": 1e3 * wall / steps,
        "nbr_ms_avg": nbr_ms / timing_steps,
        "grid_ms_avg": grid_ms / timing_steps,
        "timed_ms_per_step": total_ms / timing_steps,
        "cells": cells,
    }


def cpu_baseline(sample_agents: int):
    exe = os.path.join(REPO, "oracle", "_ref", "boids2d_f64")
    kind = "reference"
    if not os.path.exists(exe):
        exe = os.path.join(REPO, "oracle", "_build", "boids2d_port_f64")
        kind = "port"
    if not os.path.exists(exe):
        return None
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(cores))
    args = [exe] + [f"{k}={v}" for k, v in PARAMS.items()] + ["max_pos=4000.0", f"sample={sample_agents}"]
    out = subprocess.run(args, capture_output=True, text=True, check=True, env=env).stdout
    kv = dict(tok.split("=") for tok in out.split())
    secs = float(kv["sample_seconds"])
    return {"value": sample_agents / secs, "unit": "agent-steps/s", "cores": cores, "kind": kind,
            "sample": (f"update_boid for the first {sample_agents} of {N_AGENTS} agents, each "
                       f"scanning all {N_AGENTS} (the C backend's brute-force near loop, "
                       f"gcc -O2 -std=c99 -fopenmp), {secs:.2f} s"),
            "seconds": secs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--timing-steps", type=int, default=50)
    ap.add_argument("--cpu-sample", type=int, default=65536)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the 1M-agent workload fits one MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    f64 = run_gpu("f64", args.steps, args.warmup, args.timing_steps)
    f32 = None if args.no_fp32 else run_gpu("f32", args.steps, args.warmup, args.timing_steps)
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_sample)

    b = neighbour_bytes(N_AGENTS, "f64", f64["cells"])
    achieved = b / (f64["nbr_ms_avg"] * 1e-3) / 1e9
    line = {
        "metric": "agent-steps/sec, 1M boids2d FP64 vs FP32",
        "value": round(f64["agent_steps_per_s"], 1),
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f64["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (the model's own main(): xorshift128+ from the reference's fixed state)",
        "config": {"workload": "boids2d_cfg 1,048,576 agents, 4000x4000 box, radius 10, "
                               "weights 1.5/1.0/1.0, max speed 2",
                   "agents": N_AGENTS, "cells": f64["cells"], "precision": "f64",
                   "step": "grid rebuild + update_boid"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "boids_step<double,false>",
                     "bytes_per_launch": b, "avg_kernel_ms": round(f64["nbr_ms_avg"], 5)},
        "grid_ms_per_step": round(f64["grid_ms_avg"], 5),
        "cpu_baseline": cpu and {k: v for k, v in cpu.items() if k != "seconds"},
    }
    if f32:
        b32 = neighbour_bytes(N_AGENTS, "f32", f32["cells"])
        line["fp32"] = {"value": round(f32["agent_steps_per_s"], 1),
                        "ms_per_step": round(f32["ms_per_step"], 5),
                        "avg_kernel_ms": round(f32["nbr_ms_avg"], 5),
                        "grid_ms_per_step": round(f32["grid_ms_avg"], 5),
                        "kernel_GBs": round(b32 / (f32["nbr_ms_avg"] * 1e-3) / 1e9, 1)}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
