"""Benchmark: agent-steps/s of 1M-agent boids2d on one MI355X, FP64 and FP32.

Workload (1,048,576 boids, the metric's configuration): examples/boids2d_cfg.abl
with agent_density 0.065536 (a 4000 x 4000 box), positions uniform, velocities
uniform in [-1,1]^2, interaction radius 10, separation radius 2, weights
separation/alignment/cohesion 1.5/1.0/1.0, max speed 2; initial state from the
model's own main() and the reference's fixed xorshift128+ state.  A "step" is
one timestep of simulate(), agents resident in HBM: the update_boid neighbour
kernel walking per-agent neighbour lists, plus the device-gated rebuild
kernels (delta re-sort of the cell-sorted order and list build), which do
their work on the timesteps after some agent moved more than half the list
skin and return at once otherwise.

The JSON line carries
  value            agent-steps/s, FP64, whole timestep (grid + neighbour step);
  fp32             the same at use_float precision;
  fp32_compensated the same at FP32-compensated precision, with its error
                   (and FP32's) against the FP64 path over the first steps;
  roofline         the neighbour kernel: algorithmic bytes per launch
                   (N * 64 B + 4 * (cells + 1) for FP64) / its average duration,
                   timed with HIP events on the simulation's own stream; the
                   HBM bytes the counters saw (traffic) come from the committed
                   rocprofv3 passes of the same kernel sources, or are null;
  valu_roofline    VALU issue rate of the same kernel from the committed SQ
                   counters and this run's kernel duration (or null, as above);
  cpu_baseline     the reference C path (oracle/_ref: the restated generated
                   program linked against the reference's libabl.c, gcc -O2
                   -fopenmp, brute-force near loop) on a bounded sample: the
                   first M agents' update_boid against the full 1M population;
  serving          p50/p99 run latency and aggregate agent-steps/s of a
                   Poisson stream of mixed runs (1k-1M agents, 100-1000 steps);
  generated        the same workload as openabl-hip -b hip emits it
                   (boids2d_cfg.abl at 1M agents, generated step on the LDS tiles),
                   beside the hand-written kernels at the same launch configuration;
  ensemble         256 uneven boids2d members batched, against members alone;
  precision_circle3d  1M clustered circle3d: FP64, FP32-compensated and FP32
                   agent-steps/s and the per-step error against FP64;
  launch_config    the launch configuration of the runs: CONFIG below, keys
                   overridden with --set hip.KEY=VALUE, or chosen by
                   --autotune (openabl_amd.autotune on this workload's initial
                   state; its search is reported under "autotune").
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
VALU_PEAK_GWAVE = 256 * 4 * 2.4 / 4  # G wave64 VALU instructions/s
N_AGENTS = 1 << 20
DENSITY = 0.065536
PARAMS = dict(num_agents=N_AGENTS, agent_density=DENSITY, interaction_radius=10.0,
              separation_radius=2.0, steer_scale=1.0, collision_scale=1.5, match_scale=1.0,
              max_speed=2.0, time_scale=0.0005, global_scale=0.15)


# launch configuration of the runs (hip.cell_scale: cells per granularity;
# hip.tile: LDS tile shape of the neighbour kernel; hip.block: threads per
# block; hip.lds_slack: LDS records per expected halo agent; hip.nlist /
# hip.skin: neighbour lists and their skin as a fraction of the radius): the
# fastest of sweeps over tile shapes, cell scales, blocks, LDS budgets and
# skins on this workload (openabl_amd/autotune.py runs the same kind of
# search for any program)
CONFIG = {"hip.cell_scale": 2, "hip.tile": 1, "hip.block": 256, "hip.lds_slack": 1.25,
          "hip.nlist": 1, "hip.skin": 0.05}


def parse_value(v: str):
    """A launch-configuration value from the command line: int, float or text."""
    for cast in (int, float):
        try:
            return cast(v)
        except ValueError:
            pass
    return v


def run_autotune(steps: int = 8, warmup: int = 3) -> dict:
    """Time the autotuner's default space (threads per block, cell scale, tile
    shape, LDS budget) on the metric workload's initial state, FP64, keeping
    the list settings of CONFIG, and make the fastest the launch configuration."""
    from openabl_amd import MEMBERS, Simulation, load_library, record_dtype
    from openabl_amd.autotune import tune
    load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    sim = Simulation("boids2d", "f64", params=PARAMS, config=CONFIG)
    sim.init_agents()
    recs = [sim.download(record_dtype(MEMBERS["boids2d"], "f64"))]
    sim.close()
    t0 = time.perf_counter()
    base = {k: v for k, v in CONFIG.items() if k in ("hip.nlist", "hip.skin")}
    best, table = tune("boids2d", "f64", PARAMS, recs, steps=steps, warmup=warmup,
                       base_config=base)
    CONFIG.clear()
    CONFIG.update(best)
    ok = [row for row in table if row[1] is not None]
    return {"best": best, "candidates": len(table), "ran": len(ok),
            "best_ms_per_step": round(min(row[1] for row in ok) * 1e3, 4) if ok else None,
            "steps_per_candidate": steps, "seconds": round(time.perf_counter() - t0, 1)}


def progress(msg: str) -> None:
    """One line per finished leg on stderr (the JSON line is on stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def neighbour_bytes(n: int, precision: str, cells: int) -> int:
    """Algorithmic HBM bytes of one update_boid launch: every agent's pos and
    velocity read once and written once, plus the cell offsets.  Per agent:
    FP64 2 x 32 B, FP32 2 x 16 B, FP32-compensated 2 x (16 + 8) B."""
    per_agent = {"f64": 64, "f32": 32, "f32c": 48}[precision]
    return n * per_agent + 4 * (cells + 1)


def kernel_name(precision: str) -> str:
    """The neighbour kernel the runtime launches for CONFIG."""
    tile, cs = CONFIG["hip.tile"], CONFIG["hip.cell_scale"]
    if tile == 0:
        return {"f64": "boids_step<double,false>", "f32": "boids_step<float,false>",
                "f32c": "boids_step_c<false>"}[precision]
    pol = {"f64": "BoidPolicy<double>", "f32": "BoidPolicy<float>", "f32c": "BoidPolicyC"}[precision]
    tx = {1: 16, 2: 8, 3: 32}[tile]
    if CONFIG.get("hip.nlist", 1):
        # cells are widened by 1 + skin in list mode, so the list radius
        # (1 + skin) * r reaches cs cells
        rm = 1 if cs <= 1 else (2 if cs == 2 else 4)
        return f"tiled_step_list<{pol}, {tx}, 8, 1, {rm}>"
    rm = 1 if cs <= 1 else (2 if cs == 2 else 4)
    return f"tiled_step<{pol}, {tx}, 8, 1, {rm}>"


# the sources the neighbour kernel is compiled from: a committed profile
# figure is used only while these hash as they did when it was measured
KERNEL_SOURCES = ("openabl_amd/csrc/abl_tile.hpp", "openabl_amd/csrc/abl_device.hpp",
                  "openabl_amd/csrc/abl_pipeline.hpp", "openabl_amd/csrc/abl_runtime.hpp",
                  "openabl_amd/csrc/prog_boids2d.hip")


def kernel_source_sha() -> str:
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(REPO, rel), "rb") as f:
            h.update(rel.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def profile_figures(kernel: str, precision: str = "f64"):
    """The committed rocprofv3 counter figures of `kernel` under CONFIG
    (profiles/boids2d_1m_<prec>_counters.json: FETCH_SIZE/WRITE_SIZE passes
    scaled by a copy kernel of known size, and the SQ instruction counters),
    or (None, reason) when none match this kernel, launch configuration and
    kernel sources."""
    path = os.path.join(REPO, "profiles", f"boids2d_1m_{precision}_counters.json")
    if not os.path.exists(path):
        return None, "no committed counter profile"
    with open(path) as f:
        t = json.load(f)
    if t.get("kernel") != kernel or t.get("launch_config") != CONFIG:
        return None, "committed profile is of another kernel or launch configuration"
    if t.get("source_sha16") != kernel_source_sha():
        return None, "committed profile predates the current kernel sources"
    return t, f"committed profile {os.path.relpath(path, REPO)} (source sha {t['source_sha16']})"


def cpu_model() -> str | None:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def grid_cells(box: float, granularity: float, cell_scale: int) -> int:
    """Cells of the 2-D grid the runtime builds (cell edge granularity /
    cell_scale, widened by 2^-20, and by 1 + skin in neighbour-list mode)."""
    cell = granularity / cell_scale * (1.0 + 1.0 / 1048576.0)
    if CONFIG.get("hip.nlist", 1):
        cell *= 1.0 + CONFIG.get("hip.skin", 0.05)
    return max(1, math.ceil(box / cell)) ** 2


def run_gpu(precision: str, steps: int, warmup: int, timing_steps: int):
    import openabl_amd
    from openabl_amd import Simulation, load_library
    lib = load_library()
    lib.abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    sim = Simulation("boids2d", precision, params=PARAMS, config=CONFIG)
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
    cells = grid_cells(sim.get_param("max_pos"), PARAMS["interaction_radius"],
                       CONFIG["hip.cell_scale"])
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


def boids_precision_error(steps: int):
    """FP32-compensated and FP32 (use_float) boids2d on the metric workload,
    from the FP64 path's initial state, against the FP64 path (which the
    parity tests hold to the oracle within 1e-12 per step): the largest
    position and velocity differences after each step."""
    import numpy as np
    from openabl_amd import MEMBERS, Simulation, load_library, record_dtype
    dt = record_dtype(MEMBERS["boids2d"], "f64")
    dt32 = record_dtype(MEMBERS["boids2d"], "f32")
    load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    sims = {"f64": Simulation("boids2d", "f64", params=PARAMS, config=CONFIG)}
    sims["f64"].init_agents()
    s0 = sims["f64"].download(dt)
    for prec in ("f32c", "f32"):
        sims[prec] = Simulation("boids2d", prec, params=PARAMS, config=CONFIG)
        sims[prec].upload(s0 if prec == "f32c" else s0.astype(dt32))
    err = {p: {"pos_max_abs": [], "vel_max_abs": []} for p in ("f32c", "f32")}
    for _ in range(steps):
        for sim in sims.values():
            sim.simulate(1)
        want = sims["f64"].download(dt)
        for prec in ("f32c", "f32"):
            got = sims[prec].download(dt if prec == "f32c" else dt32)
            for key, col in (("pos_max_abs", "pos"), ("vel_max_abs", "velocity")):
                d = np.abs(got[col].astype(np.float64) - want[col]).max()
                err[prec][key].append(float(f"{d:.3e}"))
    for sim in sims.values():
        sim.close()
    return {"steps": steps, "reference": "the FP64 path from the same initial state",
            "box": 4000.0, "max_speed": PARAMS["max_speed"], **err}


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
            "cpu_model": cpu_model(),
            "sample": (f"update_boid for the first {sample_agents} of {N_AGENTS} agents, each "
                       f"scanning all {N_AGENTS} (the C backend's brute-force near loop, "
                       f"gcc -O2 -std=c99 -fopenmp), {secs:.2f} s"),
            "seconds": secs}


def ensemble_members(k=256, seed=2026):
    """The batched-ensemble workload: boids2d members with agent counts
    log-uniform in [1k, 64k], box side sqrt(N/0.065), weights U[0.5, 2]."""
    import math
    import numpy as np
    from openabl_amd import fold6
    rng = np.random.default_rng(seed)
    out = []
    for n in np.exp(rng.uniform(math.log(1000), math.log(65536), k)).astype(int):
        w = rng.uniform(0.5, 2.0, 3)
        out.append({"num_agents": int(n), "agent_density": 0.065,
                    "max_pos": fold6(math.sqrt(n / 0.065)), "steer_scale": float(w[0]),
                    "collision_scale": float(w[1]), "match_scale": float(w[2])})
    return out


def run_generated(steps: int, warmup: int):
    """The metric workload as `openabl-hip -b hip` emits it: examples/boids2d_cfg.abl
    generated with -P num_agents=1048576 (prebuilt by __graft_entry__.build()
    as openabl_amd/gen/bench_boids2d_cfg), its update_boid on the runtime's
    LDS tiles (the generated default), from the same initial state, FP64;
    and its results after the timed steps against the hand-written kernels'
    walking the same tiled stencil (bit for bit)."""
    from openabl_amd import MEMBERS, Simulation, codegen, load_library, record_dtype
    d = codegen.program_dir("bench_boids2d_cfg")
    if not os.path.exists(os.path.join(d, "libsteps.so")):
        return {"skipped": "bench_boids2d_cfg not built"}
    codegen.load_program(d)
    dt = record_dtype(MEMBERS["boids2d"], "f64")
    load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    s0 = Simulation("boids2d", "f64", params=PARAMS, config=CONFIG)
    s0.init_agents()
    rec = s0.download(dt)
    s0.close()
    cfg = {"hip.cell_scale": CONFIG["hip.cell_scale"], "hip.tile": CONFIG["hip.tile"],
           "hip.block": CONFIG["hip.block"], "hip.lds_slack": CONFIG["hip.lds_slack"],
           "hip.nlist": 0}
    out = {}
    for prog, params in (("bench_boids2d_cfg", None), ("boids2d", PARAMS)):
        sim = Simulation(prog, "f64", params=params, config=cfg)
        sim.upload(rec)
        sim.simulate(warmup)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.simulate(steps)
        sim.synchronize()
        wall = time.perf_counter() - t0
        out[prog] = (N_AGENTS * steps / wall, 1e3 * wall / steps, sim.download(dt).tobytes())
        sim.close()
    g, h = out["bench_boids2d_cfg"], out["boids2d"]
    return {"metric": "agent-steps/s, 1M boids2d_cfg FP64 as openabl-hip -b hip generates it",
            "value": round(g[0], 1), "ms_per_step": round(g[1], 5),
            "handwritten_same_config": round(h[0], 1),
            "results_equal_handwritten": g[2] == h[2],
            "config": dict(cfg, program="examples/boids2d_cfg.abl -P num_agents=1048576",
                           near_loop="LDS tiles, stencil walk (lists are opt-in: -C hip.lists=1)")}


def run_ensemble(args):
    """256 uneven boids2d members, one launch per kernel, graph-replayed
    timesteps through neighbour lists (each member re-sorted and rebuilt on
    its own schedule); beside it the same members run alone one after another
    (a sample of them, extrapolated) on the plain per-run path."""
    from openabl_amd import Simulation, load_library
    from openabl_amd.ensemble import Ensemble, member_seed
    ms = ensemble_members()
    total = sum(m["num_agents"] for m in ms)
    res = {}
    for prec in ("f64", "f32"):
        ens = Ensemble("boids2d", prec, ms, config={"hip.cell_scale": 2, "hip.tile": 1})
        ens.init(0xDEADBEEF, 0xBEEFDEAD)
        ens.simulate(args.warmup)
        ens.synchronize()
        t0 = time.perf_counter()
        ens.simulate(args.ens_steps)
        dev_ms = ens.synchronize()
        wall = time.perf_counter() - t0
        ens.close()
        # plain path: members alone, sequentially (every 8th member, scaled)
        sample = ms[::8]
        t_alone = 0.0
        for k, p in enumerate(sample):
            load_library().abl_rng_seed(*member_seed(0xDEADBEEF, 0xBEEFDEAD, 8 * k))
            sim = Simulation("boids2d", prec, params=p, config={"hip.cell_scale": 2, "hip.tile": 1})
            sim.init_agents()
            sim.simulate(5)
            sim.synchronize()
            t1 = time.perf_counter()
            sim.simulate(args.alone_steps)
            sim.synchronize()
            t_alone += (time.perf_counter() - t1) / args.alone_steps
            sim.close()
        sample_agents = sum(p["num_agents"] for p in sample)
        alone_rate = sample_agents / t_alone
        res[prec] = {"agent_steps_per_s": round(total * args.ens_steps / wall, 1),
                     "device_ms_per_step": round(dev_ms / args.ens_steps, 5),
                     "alone_sequential_agent_steps_per_s": round(alone_rate, 1),
                     "speedup_vs_alone": round(total * args.ens_steps / wall / alone_rate, 2)}
    return {"metric": "batched ensemble agent-steps/s (256 boids2d members)",
            "members": len(ms), "agents": total, "steps": args.ens_steps,
            "config": "N log-uniform [1k, 64k], box sqrt(N/0.065), weights U[0.5,2], "
                      "cell_scale 2, 16x8 tiles, neighbour lists (skin 0.05, rebuilds per "
                      "member), graph-replayed timesteps",
            **res}


CIRCLE3D_PARAMS = dict(num_agents=1 << 20, rho=(1 << 20) / 1e9, r=4.0, k_rep=0.05, k_att=0.01,
                       num_clusters=64, cluster_sigma=20.0, background_fraction=0.2)
# clustered 3-D: cluster cores hold ~50x the mean density, so the untiled
# one-thread-per-agent kernel (balanced by construction) beats LDS tiles,
# whose dense tiles serialise; cells of half the granularity (sweep on this
# workload)
CIRCLE3D_CONFIG = {"hip.cell_scale": 2, "hip.tile": 0, "hip.block": 256}


def run_precision(args):
    """circle3d, 1,048,576 agents in a 1000^3 box from 64 Gaussian clusters
    (sigma 20) plus 20% uniform background, radius 8: agent-steps/s of the
    FP64 path and of FP32-compensated, and the f32c error against the FP64
    path per step over 100 steps from the same initial state."""
    import numpy as np
    from openabl_amd import MEMBERS, Simulation, load_library, record_dtype
    res = {}
    for prec in ("f64", "f32c", "f32"):
        load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
        sim = Simulation("circle3d", prec, params=CIRCLE3D_PARAMS, config=CIRCLE3D_CONFIG)
        sim.init_agents()
        sim.simulate(args.warmup)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.simulate(args.prec_steps)
        sim.synchronize()
        wall = time.perf_counter() - t0
        sim.set_config("hip.timing", 1)
        sim.simulate(10)
        _, nbr_ms, grid_ms = sim.last_timing()
        sim.close()
        res[prec] = {"agent_steps_per_s": round(N_AGENTS * args.prec_steps / wall, 1),
                     "ms_per_step": round(1e3 * wall / args.prec_steps, 5),
                     "nbr_kernel_ms": round(nbr_ms / 10, 5), "grid_ms": round(grid_ms / 10, 5)}
    # error: every precision from the FP64 initial state, compared every step
    # with the FP64 path; FP32 (use_float) beside f32c for scale
    dt = record_dtype(MEMBERS["circle3d"], "f64")
    dt32 = record_dtype(MEMBERS["circle3d"], "f32")
    sims = {}
    for prec in ("f64", "f32c", "f32"):
        load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
        sims[prec] = Simulation("circle3d", prec, params=CIRCLE3D_PARAMS, config=CIRCLE3D_CONFIG)
        if prec == "f64":
            sims[prec].init_agents()
            s0 = sims[prec].download(dt)
        else:
            sims[prec].upload(s0 if prec == "f32c" else s0.astype(dt32))
    abs_err = {"f32c": [], "f32": []}
    rel_err = {"f32c": [], "f32": []}
    for t in range(args.prec_steps):
        for s in sims.values():
            s.simulate(1)
        a = sims["f64"].download(dt)["pos"]
        for prec in ("f32c", "f32"):
            b = sims[prec].download(dt if prec == "f32c" else dt32)["pos"].astype(np.float64)
            d = np.abs(a - b)
            abs_err[prec].append(float(d.max()))
            rel_err[prec].append(float((d / np.maximum(np.abs(a), 1.0)).max()))
    for s in sims.values():
        s.close()
    return {
        "metric": "agent-steps/s, 1M circle3d clustered, FP32-compensated vs FP64",
        "value": res["f32c"]["agent_steps_per_s"], "unit": "agent-steps/s",
        "f64": res["f64"], "f32c": res["f32c"], "f32": res["f32"],
        "speedup": round(res["f32c"]["agent_steps_per_s"] / res["f64"]["agent_steps_per_s"], 3),
        "error_vs_f64_path": {
            "steps": args.prec_steps,
            "note": "trajectories from the same initial state; the model is chaotic, so "
                    "any difference grows until trajectories decorrelate",
            "f32c_max_abs_per_step": [float(f"{e:.3e}") for e in abs_err["f32c"]],
            "f32c_max_rel_per_step": [float(f"{e:.3e}") for e in rel_err["f32c"]],
            "f32_max_abs_per_step": [float(f"{e:.3e}") for e in abs_err["f32"]]},
        "config": {"workload": "circle3d_clustered 1,048,576 agents, 1000^3 box, 64 clusters "
                               "sigma 20 + 20% uniform, radius 8", "steps": args.prec_steps,
                   "launch_config": CIRCLE3D_CONFIG},
    }


def run_serving(args):
    """Poisson arrivals of circle (FP64), boids2d (FP32), game_of_life and
    predator_prey runs, agent counts log-uniform 1k-1M, 100-1000 steps,
    admitted by HBM footprint; p50/p99 latency beside aggregate agent-steps/s."""
    import math
    import numpy as np
    from openabl_amd.serve import Server
    rng = np.random.default_rng(7)
    kinds = [("circle", "f64"), ("boids2d", "f32"), ("game_of_life", "f64"),
             ("predator_prey", "f64")]
    reqs = []
    t = 0.0
    for k in range(args.serve_runs):
        prog, prec = kinds[k % len(kinds)]
        n = int(np.exp(rng.uniform(math.log(1000), math.log(args.serve_max_agents))))
        steps = int(rng.integers(100, 1001))
        if prog == "circle":
            p = {"num_agents": n, "rho": 0.1024, "r": 2.5}
        elif prog == "boids2d":
            p = {"num_agents": n, "agent_density": 0.065536}
        elif prog == "game_of_life":
            p = {"num_agents": n, "alive_fraction": 0.3}
        else:
            p = {"num_agents": n, "prey_fraction": 0.4, "predator_fraction": 0.04}
        t += rng.exponential(args.serve_interval)
        reqs.append((t, prog, prec, p, steps))
    srv = Server(hbm_budget=int(args.serve_budget_gb * (1 << 30)), max_active=args.serve_active,
                 chunk_steps=args.serve_chunk)
    t0 = time.perf_counter()
    ids = []
    for (at, prog, prec, p, steps) in reqs:
        while time.perf_counter() - t0 < at:
            time.sleep(0.0005)
        ids.append(srv.submit(prog, prec, p, steps))
    for rid in ids:
        srv.wait(rid)
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    return {"metric": "serving: p50/p99 run latency and agent-steps/s under load",
            "runs": len(reqs), "p50_s": round(st["p50_s"], 4),
            "p99_s": round(st["p99_s"], 4),
            "agent_steps_per_s": round(st["agent_steps"] / wall, 1),
            "wall_s": round(wall, 3), "peak_admitted_bytes": st["peak_bytes"],
            "config": {"mix": "circle f64, boids2d f32, game_of_life, predator_prey",
                       "agents": f"log-uniform 1k-{args.serve_max_agents}",
                       "steps": "uniform 100-1000",
                       "mean_interarrival_s": args.serve_interval,
                       "budget_gb": args.serve_budget_gb,
                       "max_active": args.serve_active, "chunk_steps": args.serve_chunk}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--timing-steps", type=int, default=50)
    ap.add_argument("--cpu-sample", type=int, default=262144)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-generated", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--error-steps", type=int, default=20)
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the runs, copy 64 MiB device to device 5 times (the HIP "
                         "runtime's copy kernel: reference for scaling HBM byte counters)")
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="override a key of the launch configuration (hip.*), e.g. hip.skin=0.04")
    ap.add_argument("--autotune", action="store_true",
                    help="choose the launch configuration with openabl_amd.autotune on this "
                         "workload before the timed runs")
    ap.add_argument("--mode", choices=["main", "ensemble", "serving", "precision"], default="main")
    ap.add_argument("--prec-steps", type=int, default=100)
    ap.add_argument("--ens-steps", type=int, default=500)
    ap.add_argument("--alone-steps", type=int, default=50)
    ap.add_argument("--serve-runs", type=int, default=40)
    ap.add_argument("--serve-max-agents", type=int, default=1 << 20)
    ap.add_argument("--serve-interval", type=float, default=0.05)
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--serve-active", type=int, default=8)
    ap.add_argument("--serve-chunk", type=int, default=25,
                    help="timesteps a served run is given at a time")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the 1M-agent workload fits one MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.mode == "ensemble":
        print(json.dumps(run_ensemble(args)))
        return
    if args.mode == "serving":
        print(json.dumps(run_serving(args)))
        return
    if args.mode == "precision":
        print(json.dumps(run_precision(args)))
        return

    for kv in args.set:
        k, _, v = kv.partition("=")
        CONFIG[k] = parse_value(v)
    tuned = run_autotune() if args.autotune else None
    if tuned:
        progress(f"autotune: {tuned['best']} of {tuned['candidates']} in {tuned['seconds']} s")
    f64 = run_gpu("f64", args.steps, args.warmup, args.timing_steps)
    progress(f"f64: {f64['agent_steps_per_s']:.4g} agent-steps/s")
    f32 = None if args.no_fp32 else run_gpu("f32", args.steps, args.warmup, args.timing_steps)
    f32c = None if args.no_fp32 else run_gpu("f32c", args.steps, args.warmup, args.timing_steps)
    if f32c:
        progress(f"f32: {f32['agent_steps_per_s']:.4g}, f32c: {f32c['agent_steps_per_s']:.4g}")
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_sample)
    if cpu:
        progress(f"cpu baseline: {cpu['value']:.4g} agent-steps/s")

    b = neighbour_bytes(N_AGENTS, "f64", f64["cells"])
    achieved = b / (f64["nbr_ms_avg"] * 1e-3) / 1e9
    kernel = kernel_name("f64")
    figures, source = profile_figures(kernel)
    valu = None
    if figures and "sq" in figures:
        # VALU issue: wave64 instructions per launch over the kernel's
        # duration, against 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64
        # instruction (FP64 and non-packed FP32 issue at that rate on MI355X)
        sq = figures["sq"]
        rate = sq["SQ_INSTS_VALU"] / (f64["nbr_ms_avg"] * 1e-3) / 1e9
        peak = VALU_PEAK_GWAVE
        valu = {"bound": "valu", "achieved": round(rate, 1), "peak": peak,
                "unit": "G wave64-instructions/s", "frac": round(rate / peak, 4),
                "lanes_active": round(sq["SQ_THREAD_CYCLES_VALU"] /
                                      (sq["SQ_ACTIVE_INST_VALU"] * 64.0), 3),
                "valu_per_wave": round(sq["SQ_INSTS_VALU"] / sq["SQ_WAVES"], 1),
                "source": source}
    line = {
        "metric": "agent-steps/sec, 1M boids2d FP64 vs FP32",
        "launch_config": CONFIG,
        "autotune": tuned,
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
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": figures and figures["traffic_bytes"],
                     "traffic_source": source,
                     "kernel": kernel,
                     "bytes_per_launch": b, "avg_kernel_ms": round(f64["nbr_ms_avg"], 5)},
        "valu_roofline": valu or {"source": source},
        "grid_ms_per_step": round(f64["grid_ms_avg"], 5),
        "cpu_baseline": cpu and {k: v for k, v in cpu.items() if k != "seconds"},
    }
    for key, prec, r in (("fp32", "f32", f32), ("fp32_compensated", "f32c", f32c)):
        if r:
            bp = neighbour_bytes(N_AGENTS, prec, r["cells"])
            line[key] = {"value": round(r["agent_steps_per_s"], 1),
                         "ms_per_step": round(r["ms_per_step"], 5),
                         "avg_kernel_ms": round(r["nbr_ms_avg"], 5),
                         "grid_ms_per_step": round(r["grid_ms_avg"], 5),
                         "kernel_GBs": round(bp / (r["nbr_ms_avg"] * 1e-3) / 1e9, 1)}
    if f32c and args.error_steps > 0:
        line["fp32_compensated"]["error_vs_f64"] = boids_precision_error(args.error_steps)
        progress("f32c error leg done")
    if not args.no_generated:
        line["generated"] = run_generated(min(args.steps, 100), args.warmup)
        progress("generated leg done")
    if not args.no_ensemble:
        line["ensemble"] = run_ensemble(args)
        progress("ensemble leg done")
    if not args.no_precision:
        line["precision_circle3d"] = run_precision(args)
        progress("circle3d precision leg done")
    if not args.no_serving:
        line["serving"] = run_serving(args)
        progress("serving leg done")
    print(json.dumps(line))
    if args.copy_probe:
        import torch
        x = torch.empty(8 << 20, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
