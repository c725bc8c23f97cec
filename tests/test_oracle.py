"""CPU tests that pin the oracle before anything is compared with it.

The oracle (oracle/) restates the reference C backend's generated programs
and its runtime asset/c/libabl.{h,c}.  It is pinned by:
  * golden fixtures written from the same programs linked against the
    reference's own libabl.c (tests/golden/make_golden.py, oracle/_ref),
  * a live comparison with oracle/_ref where it has been built,
  * closed forms: Conway's B3/S23 rule on a bounded grid (game_of_life), the
    reference's near test recomputed with numpy's correctly rounded float32
    sqrt, and the grid path agreeing bit-for-bit with the brute-force loop.
"""
import hashlib
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from orclib import (Oracle, field_bytes, boids2d_params, circle3d_params, circle_params, game_of_life_params,
                    ref_binary, BUILD_DIR, REF_DIR)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def probe(prec):
    return os.path.join(BUILD_DIR, f"orc_probe_{prec}")


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_rng_matches_reference_draws(prec):
    """xorshift128+ / random_float / random_int (asset/c/libabl.c:8-39),
    including the power-of-two random_int branch that skips `min`."""
    for case in load_golden(f"rng_{prec}.json"):
        if case["kind"] == "mix":
            out = subprocess.run([probe(prec), "mix", "40"], capture_output=True, text=True,
                                 check=True).stdout.split("\n")[:-1]
        else:
            out = subprocess.run([probe(prec), case["kind"], str(case["count"]), str(case["lo"]),
                                  str(case["hi"])], capture_output=True, text=True,
                                 check=True).stdout.split()
        assert out == case["values"], case["kind"]


def test_rng_first_draws_closed_form():
    """Independent restatement of xorshift128+ in Python integers."""
    s0, s1 = 0xDEADBEEF, 0xBEEFDEAD
    M = (1 << 64) - 1
    vals = []
    for _ in range(8):
        x, y = s0, s1
        s0 = y
        x = (x ^ (x << 23)) & M
        s1 = x ^ y ^ (x >> 17) ^ (y >> 26)
        vals.append(((s1 + y) & M) / 2.0 ** 64)
    golden = [float.fromhex(v) for v in load_golden("rng_f64.json")[0]["values"][:8]]
    assert vals == golden
    # first draws quoted by the survey of the reference: 0.001699..., 0.002433...
    assert abs(vals[0] - 0.0016989) < 1e-6 and abs(vals[1] - 0.0024333) < 1e-6


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_save_json_matches_reference(prec, tmp_path):
    """save() text (asset/c/libabl.c:46-124): "%f", vectors, bools, ",\\n"."""
    p = tmp_path / "s.json"
    subprocess.run([probe(prec), "save", str(p)], check=True)
    assert p.read_text() == load_golden(f"save_{prec}.json")["text"]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_dist_matches_reference(prec, tmp_path):
    """dist_float2/3 = (abl_float) sqrtf((float) sum of squares) on pairs at
    radius +- 1e-9..1e-5 relative (asset/c/libabl.h:156-168)."""
    g = load_golden(f"dist_{prec}.json")
    for dim in (2, 3):
        case = g[f"dim{dim}"]
        p = tmp_path / f"p{dim}.txt"
        p.write_text("\n".join(" ".join(row) for row in case["pairs"]) + "\n")
        out = subprocess.run([probe(prec), f"dist{dim}", str(p)], capture_output=True, text=True,
                             check=True).stdout.split()
        assert out == case["dist"]
        # numpy restatement: float32 sqrt is correctly rounded (IEEE-754)
        pairs = np.array([[float.fromhex(v) for v in row] for row in case["pairs"]])
        a, b = pairs[:, :dim], pairs[:, dim:]
        if prec == "f32":
            a, b = a.astype(np.float32), b.astype(np.float32)
        d = a - b
        s = (d * d).sum(axis=1) if dim == 2 else (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        want = np.sqrt(s.astype(np.float32)).astype(np.float64)
        assert np.array_equal(want, np.array([float.fromhex(v) for v in out]))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_model_runs_match_reference_runtime(prec, tmp_path):
    """The restated programs through the oracle runtime write the same bytes
    as the same programs linked against the reference's libabl.c."""
    for case in load_golden(f"runs_{prec}.json"):
        o = Oracle(case["model"], prec, grid=True).set(**case["params"]).init()
        o.simulate(int(case["params"]["num_timesteps"]))
        out = tmp_path / "o.json"
        o.save(str(out))
        data = out.read_bytes()
        assert hashlib.sha256(data).hexdigest() == case["sha256"], case["model"]
        o.close()


@pytest.mark.parametrize("model,params", [
    ("circle", circle_params(num_agents=1024)),
    ("boids2d", boids2d_params(num_agents=4096)),
    ("circle3d", circle3d_params(num_agents=3000, num_clusters=4, cluster_sigma=3.0)),
    ("game_of_life", game_of_life_params(num_agents=48 * 48, alive_fraction=0.3)),
])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_grid_oracle_equals_brute_force(model, params, prec):
    """The grid path visits the brute-force loop's candidates in the same
    order: neighbour sets and states after 5 steps are bit-identical."""
    a = Oracle(model, prec, grid=False).set(**params).init()
    b = Oracle(model, prec, grid=True).set(**params).init()
    ca, na = a.neighbours(0)
    cb, nb = b.neighbours(0)
    assert np.array_equal(ca, cb) and np.array_equal(na, nb)
    a.simulate(5)
    b.simulate(5)
    assert field_bytes(a.agents()) == field_bytes(b.agents())


def test_near_test_closed_form():
    """Neighbour sets of circle equal the numpy restatement of the C test
    (double sum of squares -> float -> correctly rounded sqrtf -> <= r),
    computed by brute force over all pairs."""
    o = Oracle("circle", "f64").set(**circle_params(num_agents=600, rho=0.06)).init()
    pos = o.agents()["pos"]
    c, n = o.neighbours(0)
    d = pos[None, :, :] - pos[:, None, :]  # nx.pos - in.pos
    s = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]
    dist = np.sqrt(s.astype(np.float32)).astype(np.float64)
    inc = ~(dist > 2.0 * o.get("r"))
    ci, ni = np.nonzero(inc)
    assert np.array_equal(c, ci) and np.array_equal(n, ni)
    # self is always a neighbour (distance 0 <= r)
    assert np.all(inc[np.arange(len(pos)), np.arange(len(pos))])


def conway(alive: np.ndarray) -> np.ndarray:
    """B3/S23 on a bounded, non-wrapping grid, 8-neighbourhood."""
    p = np.pad(alive.astype(np.int32), 1)
    n = sum(p[1 + dx:p.shape[0] - 1 + dx, 1 + dy:p.shape[1] - 1 + dy]
            for dx in (-1, 0, 1) for dy in (-1, 0, 1) if (dx, dy) != (0, 0))
    return (n == 3) | (alive & (n == 2))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_game_of_life_is_conway(prec):
    """Radius 1.5 admits sqrt(2) and excludes 2: the model is Conway's life."""
    side = 40
    o = Oracle("game_of_life", prec).set(**game_of_life_params(num_agents=side * side,
                                                               alive_fraction=0.35)).init()
    grid = o.agents()["alive"].astype(bool).reshape(side, side)  # x-major, y-minor
    for _ in range(12):
        o.simulate(1)
        grid = conway(grid)
        assert np.array_equal(o.agents()["alive"].astype(bool).reshape(side, side), grid)


def test_game_of_life_debug_blinker():
    """debug=true seeds a blinker at x = 10..12, y = 10 (period 2)."""
    o = Oracle("game_of_life").set(**game_of_life_params(num_agents=400, debug=True)).init()
    a0 = o.agents()["alive"].copy()
    assert a0.sum() == 3
    o.simulate(1)
    a1 = o.agents()["alive"].copy()
    assert a1.sum() == 3 and not np.array_equal(a0, a1)
    o.simulate(1)
    assert np.array_equal(o.agents()["alive"], a0)


def test_circle_init_uses_gcc_argument_order():
    """random(float2(W)) draws y first (gcc evaluates float2_create's
    arguments right to left): agent 0 is (draw 2, draw 1) * W."""
    g = [float.fromhex(v) for v in load_golden("rng_f64.json")[0]["values"][:2]]
    o = Oracle("circle").set(**circle_params(num_agents=4)).init()
    W = o.get("W")
    p0 = o.agents()["pos"][0]
    assert p0[0] == 0.0 + (2.0 ** 64 * g[1]) / (2.0 ** 64 / W)
    assert p0[1] == 0.0 + (2.0 ** 64 * g[0]) / (2.0 ** 64 / W)


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_DIR, "circle_f64")),
                    reason="oracle/_ref not built (needs the reference sources)")
def test_live_reference_build_matches_oracle(tmp_path):
    """Fresh comparison with the reference-runtime build, other parameters."""
    params = boids2d_params(num_agents=1500, num_timesteps=7, steer_scale=0.65,
                            collision_scale=0.75, match_scale=1.25, max_speed=1.0)
    out = tmp_path / "r.json"
    subprocess.run([ref_binary("boids2d")] + [f"{k}={v}" for k, v in params.items()]
                   + [f"out={out}"], check=True, capture_output=True)
    o = Oracle("boids2d").set(**params).init()
    o.simulate(7)
    o.save(str(tmp_path / "o.json"))
    assert out.read_bytes() == (tmp_path / "o.json").read_bytes()
