"""Driving the -b hip code generator (openabl_amd/bin/openabl-hip) from Python.

`generate()` runs the compiler on a model into a directory, `build()` runs
that directory's build.sh (hipcc for the device program, g++ for the host
program), `load_program()` loads a built step library into this process so
that its program can be simulated through openabl_amd.Simulation, and
`run_main()` runs the generated host program the way the reference's run.sh
runs its program.

PROGRAMS lists the generated programs `__graft_entry__.build()` prebuilds
under openabl_amd/gen/ (they travel with the tree, so nothing is built on a
GPU box): the example models the benchmark configurations and the parity
tests use, at the parameters of the reference-runtime fixtures
(tests/golden/runs_*.json) and in both float precisions.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess

from .runtime import load_library

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
ABLC = os.path.join(PKG, "bin", "openabl-hip")
ASSET = os.path.join(PKG, "asset")
GEN = os.path.join(PKG, "gen")
OWN_EXAMPLES = os.path.join(REPO, "examples")
REF_EXAMPLES = "/root/reference/examples"


def model_path(name: str) -> str | None:
    """The .abl file of a model: the repository's examples/ first (the
    benchmark variants), then the reference checkout's, when present."""
    for d in (OWN_EXAMPLES, REF_EXAMPLES):
        p = os.path.join(d, name + ".abl")
        if os.path.exists(p):
            return p
    return None


def generate(model: str, out_dir: str, params: dict | None = None, config: dict | None = None,
             check: bool = True) -> subprocess.CompletedProcess:
    # the asset directory compiled into openabl-hip (this tree's asset/)
    args = [ABLC, "-i", model, "-o", out_dir, "-b", "hip"]
    for k, v in (params or {}).items():
        args += ["-P", f"{k}={v}"]
    for k, v in (config or {}).items():
        args += ["-C", f"{k}={v}"]
    r = subprocess.run(args, capture_output=True, text=True)
    if check and r.returncode != 0:
        raise RuntimeError(f"openabl-hip failed ({r.returncode}): {r.stderr}")
    return r


def build(out_dir: str) -> None:
    r = subprocess.run(["./build.sh"], cwd=out_dir, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build.sh failed in {out_dir}:\n{r.stderr[-4000:]}")


_loaded: dict[str, C.CDLL] = {}


def load_program(out_dir: str) -> C.CDLL:
    """Load a generated step library (it registers its program with the
    runtime loaded by openabl_amd.load_library)."""
    load_library()
    path = os.path.join(out_dir, "libsteps.so")
    if path not in _loaded:
        if not os.path.exists(path):
            raise ImportError(f"{path} is not built; run __graft_entry__.build()")
        _loaded[path] = C.CDLL(path)
    return _loaded[path]


def run_main(out_dir: str, cwd: str, config: dict | None = None,
             timeout: float = 600) -> subprocess.CompletedProcess:
    """Run the generated host program in `cwd` (where it writes its save()
    files); `config` becomes ABL_HIP_CONFIG (run-time -C hip.* keys)."""
    env = dict(os.environ)
    if config:
        env["ABL_HIP_CONFIG"] = ",".join(f"{k}={v}" for k, v in config.items())
    return subprocess.run([os.path.join(out_dir, "main")], cwd=cwd, env=env, capture_output=True,
                          text=True, timeout=timeout)


# name -> (model, -P parameters, -C configuration); the program registers as
# `name`.  The *_golden_* programs use the parameters of the reference-runtime
# fixtures in tests/golden/runs_<prec>.json.
def _programs():
    golden = {
        "circle": ("circle", {"num_agents": 1024, "rho": 0.1024, "r": 2.5, "num_timesteps": 100}),
        "boids2d": ("boids2d_cfg", {"num_agents": 2048, "agent_density": 0.065536,
                                    "num_timesteps": 20}),
        "circle3d": ("circle3d_clustered", {"num_agents": 2000, "num_clusters": 0,
                                            "num_timesteps": 10}),
        "circle3d_c": ("circle3d_clustered", {"num_agents": 4096, "rho": 4096 / 1e6, "r": 4.0,
                                              "num_clusters": 8, "cluster_sigma": 20.0,
                                              "num_timesteps": 5}),
        "game_of_life": ("game_of_life", {"num_agents": 4096, "alive_fraction": 0.3,
                                          "num_timesteps": 100}),
    }
    out = {}
    for case, (m, p) in golden.items():
        for prec in ("f64", "f32"):
            cfg = {"use_float": "true"} if prec == "f32" else {}
            out[f"golden_{case}_{prec}"] = (m, p, cfg)
    # models at their own defaults, for the GPU parity of the generated steps
    for m in ("circle", "circle3d", "boids2d", "boids2d_cfg", "circle3d_clustered",
              "game_of_life", "predator_prey", "boids", "boids2d_flockers", "ants", "sugarscape",
              "keratinocyte"):
        out[f"model_{m}"] = (m, {}, {})
    # boids2d_cfg with its near loop on the runtime's tiled neighbour lists
    out["model_boids2d_lists"] = ("boids2d_cfg", {}, {"hip.lists": "1"})
    # the metric workload as the generator emits it (bench.py's generated leg)
    out["bench_boids2d_cfg"] = ("boids2d_cfg", {"num_agents": 1048576}, {})
    return out


PROGRAMS = _programs()


def program_dir(name: str) -> str:
    return os.path.join(GEN, name)


def manifest() -> dict:
    p = os.path.join(GEN, "manifest.json")
    if not os.path.exists(p):
        return {}
    with open(p) as f:
        return json.load(f)


def build_all(jobs: int = 8) -> dict:
    """Generate and build every program of PROGRAMS whose model is present;
    record the outcome of each in openabl_amd/gen/manifest.json."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(GEN, exist_ok=True)
    result = {}

    def one(name):
        model, params, cfg = PROGRAMS[name]
        path = model_path(model)
        if not path:
            return name, {"status": "no model", "model": model}
        d = program_dir(name)
        r = generate(path, d, params, dict(cfg, **{"hip.program": name}), check=False)
        if r.returncode == 2:
            return name, {"status": "unsupported", "model": model, "message": r.stderr.strip()}
        if r.returncode != 0:
            raise RuntimeError(f"openabl-hip failed on {name}: {r.stderr}")
        build(d)
        return name, {"status": "built", "model": model, "params": params, "config": cfg}

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for name, info in ex.map(one, sorted(PROGRAMS)):
            result[name] = info
    with open(os.path.join(GEN, "manifest.json"), "w") as f:
        json.dump(result, f, indent=1, sort_keys=True)
    return result
