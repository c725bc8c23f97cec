"""Write the golden fixtures under tests/golden/ from the reference-runtime
builds in oracle/_ref (make -C oracle ref).

Those builds are the restated C-backend programs and a small probe, compiled
against the reference's own asset/c/libabl.c with the reference's gcc flags
(src/backend/CBackend.cpp:21-27).  The fixtures pin the oracle and the
product's host runtime without the reference being present:

  rng_<prec>.json         random_float / random_int draw sequences (hex floats)
  save_<prec>.json        save() output of a fixed two-type record set
  save_xml_<prec>.json    the same kind of records saved as JSON and as FLAME /
                          FlameGPU XML
  dist_<prec>.json        dist_float2 / dist_float3 on pairs near rounding edges
  <model>_<prec>.json     final save() text of small runs (sha256 + agent count)

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref")
sys.path.insert(0, os.path.join(REPO, "tests"))
from orclib import boids2d_params, circle3d_params, circle_params, game_of_life_params  # noqa: E402

RNG_CASES = [("float", 256, 0, 1), ("float", 64, -1, 1), ("float", 64, 0, 100),
             ("float", 64, 0, 1000), ("int", 128, 0, 9), ("int", 128, 0, 7), ("int", 64, 5, 12),
             ("int", 64, -3, 4)]

RUNS = {
    "circle": [dict(circle_params(num_agents=1024, num_timesteps=100))],
    "boids2d": [dict(boids2d_params(num_agents=2048, num_timesteps=20))],
    "circle3d": [dict(circle3d_params(num_agents=2000, num_timesteps=10)),
                 dict(circle3d_params(num_agents=4096, rho=4096 / 1e6, r=4.0, num_clusters=8,
                                      cluster_sigma=20.0, num_timesteps=5))],
    "game_of_life": [dict(game_of_life_params(num_agents=64 * 64, alive_fraction=0.3,
                                              num_timesteps=100))],
}


def run(cmd):
    return subprocess.run(cmd, check=True, capture_output=True, text=True).stdout


def dist_pairs(dim: int, seed: int, n: int = 400) -> np.ndarray:
    """Pairs whose squared distance sits next to float rounding boundaries
    and next to common radii."""
    rng = np.random.default_rng(seed)
    out = []
    radii = [1.5, 2.0, 4.0, 5.0, 8.0, 10.0]
    for k in range(n):
        a = rng.uniform(0, 1000, dim)
        r = radii[k % len(radii)]
        d = rng.normal(size=dim)
        d /= np.linalg.norm(d)
        scale = r * (1.0 + rng.choice([-1, 1]) * 10.0 ** rng.uniform(-9, -5))
        b = a + d * scale
        out.append(np.concatenate([a, b]))
    return np.array(out)


def main():
    if not os.path.exists(os.path.join(REF, "ref_probe_f64")):
        sys.exit("oracle/_ref is not built: make -C oracle ref")
    for prec in ("f64", "f32"):
        probe = os.path.join(REF, f"ref_probe_{prec}")
        rng = [{"kind": k, "count": c, "lo": lo, "hi": hi,
                "values": run([probe, k, str(c), str(lo), str(hi)]).split()}
               for (k, c, lo, hi) in RNG_CASES]
        rng.append({"kind": "mix", "count": 40, "values": run([probe, "mix", "40"]).split("\n")[:-1]})
        with open(os.path.join(HERE, f"rng_{prec}.json"), "w") as f:
            json.dump(rng, f, indent=0)
        with tempfile.TemporaryDirectory() as td:
            p = os.path.join(td, "s.json")
            run([probe, "save", p])
            with open(p) as f:
                text = f.read()
        with open(os.path.join(HERE, f"save_{prec}.json"), "w") as f:
            json.dump({"text": text}, f)
        # the same records as JSON and as FLAME / FlameGPU XML (save() with
        # SAVE_FLAME_XML, SAVE_FLAMEGPU_XML)
        xml = {}
        for kind in ("flame", "flamegpu"):
            with tempfile.TemporaryDirectory() as td:
                pj, px = os.path.join(td, "s.json"), os.path.join(td, "s.xml")
                run([probe, f"save_{kind}", pj, px])
                with open(pj) as f:
                    xml["json"] = f.read()
                with open(px) as f:
                    xml[kind] = f.read()
        with open(os.path.join(HERE, f"save_xml_{prec}.json"), "w") as f:
            json.dump(xml, f)
        dist = {}
        for dim in (2, 3):
            pairs = dist_pairs(dim, 7 + dim)
            with tempfile.TemporaryDirectory() as td:
                p = os.path.join(td, "pairs.txt")
                with open(p, "w") as f:
                    for row in pairs:
                        f.write(" ".join(float(v).hex() for v in row) + "\n")
                vals = run([probe, f"dist{dim}", p]).split()
            dist[f"dim{dim}"] = {"pairs": [[float(v).hex() for v in row] for row in pairs],
                                 "dist": vals}
        with open(os.path.join(HERE, f"dist_{prec}.json"), "w") as f:
            json.dump(dist, f)
        runs = []
        for model, cases in RUNS.items():
            for params in cases:
                with tempfile.TemporaryDirectory() as td:
                    out = os.path.join(td, "o.json")
                    args = [f"{k}={v}" for k, v in params.items()] + [f"out={out}"]
                    stdout = run([os.path.join(REF, f"{model}_{prec}")] + args)
                    with open(out, "rb") as f:
                        data = f.read()
                runs.append({"model": model, "params": params,
                             "sha256": hashlib.sha256(data).hexdigest(),
                             "bytes": len(data), "head": data[:400].decode(),
                             "stdout": stdout.split()[-1]})
        with open(os.path.join(HERE, f"runs_{prec}.json"), "w") as f:
            json.dump(runs, f, indent=1)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
