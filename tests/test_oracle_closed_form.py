"""Closed-form checks of the oracle's step bodies (CPU).

The oracle's model programs (oracle/models/*.c) restate what the reference
C backend prints for each example.  Fixtures produced by running them pin
nothing about the step bodies themselves, so each body is checked here on
tiny hand-built populations whose outcome follows from the model's text by
hand: the expected numbers below are worked out from the .abl source
(cited per test), not recomputed with the oracle's formula.  Each case is
chosen so that a plausible restatement slip -- a swapped weight, a wrong
sign, the agent itself left out of its own near loop (the C loop includes
it, src/backend/CPrinter.cpp:161-171), a strict test where the model has an
inclusive one, an else-if where the model has two ifs -- changes the
outcome.
"""
import math

import numpy as np
import pytest

from orclib import Oracle, boids2d_params, circle3d_params, circle_params, predator_prey_params

TOL = 1e-12


def close(a, b, tol=TOL):
    return abs(float(a) - float(b)) <= tol * max(1.0, abs(float(b)))


def one_step(model, params, recs, t=0, step=None, **kw):
    o = Oracle(model, "f64").set(**params)
    arr = np.zeros(len(recs), dtype=o.dtypes[t])
    for i, r in enumerate(recs):
        for k, v in r.items():
            arr[i][k] = v
    o.set_agents(arr, t)
    if step is None:
        o.simulate(1)
    else:
        o.run_step(step)
    out = o.agents(t)
    o.close()
    return out


# ---- circle (examples/circle.abl:41-57): r = 2.5, near radius 2r = 5,
# k_rep = 0.05, k_att = 0.01, clamp to [1, W - 1] with W = 100 -------------

CIRCLE = circle_params(num_agents=2, rho=2 / 10000.0)  # W = 100


def test_circle_params_fold_w():
    assert CIRCLE["W"] == 100.0 and CIRCLE["r"] == 2.5


def test_circle_repelling_pair():
    # d = 1 < r: each is pushed away by k_rep * d = 0.05 along the axis
    out = one_step("circle", CIRCLE, [{"pos": (50.0, 50.0)}, {"pos": (51.0, 50.0)}])
    assert close(out["pos"][0][0], 49.95) and close(out["pos"][1][0], 51.05)
    assert out["pos"][0][1] == 50.0 and out["pos"][1][1] == 50.0


def test_circle_attracting_pair():
    # r < d = 4 <= 2r: each moves k_att * (2r - d) = 0.01 towards the other
    out = one_step("circle", CIRCLE, [{"pos": (50.0, 50.0)}, {"pos": (54.0, 50.0)}])
    assert close(out["pos"][0][0], 50.01) and close(out["pos"][1][0], 53.99)


def test_circle_out_of_range_and_stacked():
    # d = 5.5 > 2r: no force; a stacked pair (d = 0) is skipped by pos_dist == 0
    out = one_step("circle", CIRCLE, [{"pos": (20.0, 20.0)}, {"pos": (25.5, 20.0)},
                                      {"pos": (70.0, 70.0)}, {"pos": (70.0, 70.0)}])
    assert [tuple(p) for p in out["pos"]] == [(20.0, 20.0), (25.5, 20.0), (70.0, 70.0),
                                             (70.0, 70.0)]


def test_circle_exact_outer_radius_included():
    # d = 2r exactly: inside the near loop (dist > r skips, so = r is kept),
    # but the attraction k_att * (2r - d) is zero: positions unchanged; at
    # d = 4.5 the same pair attracts by 0.005 each
    out = one_step("circle", CIRCLE, [{"pos": (30.0, 30.0)}, {"pos": (30.0, 35.0)},
                                      {"pos": (60.0, 60.0)}, {"pos": (60.0, 64.5)}])
    assert out["pos"][1][1] == 35.0 and out["pos"][0][1] == 30.0
    assert close(out["pos"][2][1], 60.005) and close(out["pos"][3][1], 64.495)


def test_circle_forces_superpose():
    # q1 - p - q2 on a line at spacing 1: p feels equal and opposite pushes
    # (stays); q1 is pushed by p (d = 1: 0.05) and by q2 (d = 2 < r: 0.1)
    out = one_step("circle", CIRCLE, [{"pos": (49.0, 50.0)}, {"pos": (50.0, 50.0)},
                                      {"pos": (51.0, 50.0)}])
    assert close(out["pos"][1][0], 50.0)
    assert close(out["pos"][0][0], 48.85) and close(out["pos"][2][0], 51.15)


def test_circle_clamp():
    # a lone agent outside [1, W - 1] is clamped to it
    out = one_step("circle", CIRCLE, [{"pos": (0.5, 99.5)}])
    assert tuple(out["pos"][0]) == (1.0, 99.0)


def test_circle3d_repelling_pair_along_z():
    p = circle3d_params(num_agents=2, rho=2 / 1e6)  # W = 100, r = 5: near radius 10
    out = one_step("circle3d", p, [{"pos": (50.0, 50.0, 50.0)}, {"pos": (50.0, 50.0, 52.0)}])
    # d = 2 < r: pushed apart by k_rep * 2 = 0.1 each, only along z
    assert close(out["pos"][0][2], 49.9) and close(out["pos"][1][2], 52.1)
    assert tuple(out["pos"][0][:2]) == (50.0, 50.0)


def test_circle3d_attracting_pair_diagonal():
    p = circle3d_params(num_agents=2, rho=2 / 1e6)
    # d = 6*sqrt(3) ~ 10.39 > 2r: no force; d = 4*sqrt(3) ~ 6.93 in (r, 2r]:
    # attraction of magnitude k_att * (10 - d) split equally over x, y, z
    d = 4.0 * math.sqrt(3.0)
    out = one_step("circle3d", p, [{"pos": (20.0, 20.0, 20.0)}, {"pos": (26.0, 26.0, 26.0)},
                                   {"pos": (60.0, 60.0, 60.0)}, {"pos": (64.0, 64.0, 64.0)}])
    assert tuple(out["pos"][0]) == (20.0, 20.0, 20.0)
    step = 0.01 * (10.0 - d) / math.sqrt(3.0)
    for k in range(3):
        assert close(out["pos"][2][k], 60.0 + step, 1e-6)  # sqrtf rounds d to float
        assert close(out["pos"][3][k], 64.0 - step, 1e-6)


# ---- boids2d_cfg (examples/boids2d.abl:44-101 with its weights as params) --
# distinct weights so that any two swapped change the outcome

BOIDS = boids2d_params(num_agents=2, agent_density=2 / 1e6, steer_scale=0.65,
                       collision_scale=0.75, match_scale=1.25, max_speed=2.0)


def test_boids_params():
    assert BOIDS["max_pos"] == 1000.0 and BOIDS["interaction_radius"] == 10.0


def test_boids_isolated():
    # alone, the boid is its own neighbour (count 1, centre = its position):
    # steer and avoid vanish, match gives v * 1.25; v' = v (1 + 1.25 * 0.15)
    out = one_step("boids2d", BOIDS, [{"pos": (500.0, 500.0), "velocity": (1.0, 0.5)}])
    assert close(out["velocity"][0][0], 1.1875) and close(out["velocity"][0][1], 0.59375)
    assert close(out["pos"][0][0], 500.0 + 1.1875 * 0.0005)
    assert close(out["pos"][0][1], 500.0 + 0.59375 * 0.0005)


def test_boids_speed_clamp():
    # v' = (3, 4) * 1.1875 has length 5.9375 > 2: scaled to length 2, same direction
    out = one_step("boids2d", BOIDS, [{"pos": (500.0, 500.0), "velocity": (3.0, 4.0)}])
    assert close(out["velocity"][0][0], 1.2) and close(out["velocity"][0][1], 1.6)


def test_boids_pair_inside_separation():
    # d = 1 < 2: both neighbours (itself and the other) within both radii, so
    # the centre and the collision centre are both the midpoint (0.5 away):
    # v'.x = 0.15 * 0.5 * (steer - collision) = -0.0075 for the left boid
    out = one_step("boids2d", BOIDS, [{"pos": (500.0, 500.0), "velocity": (0.0, 0.0)},
                                      {"pos": (501.0, 500.0), "velocity": (0.0, 0.0)}])
    assert close(out["velocity"][0][0], -0.0075) and close(out["velocity"][1][0], 0.0075)
    assert out["velocity"][0][1] == 0.0


def test_boids_pair_cohesion_and_match():
    # d = 5: cohesion towards the midpoint 2.5 away (2.5 * 0.65), collision
    # only with itself (no avoidance), match of the mean velocity
    # ((1,0) + (0,1)) / 2 * 1.25 = (0.625, 0.625)
    out = one_step("boids2d", BOIDS, [{"pos": (500.0, 500.0), "velocity": (1.0, 0.0)},
                                      {"pos": (505.0, 500.0), "velocity": (0.0, 1.0)}])
    assert close(out["velocity"][0][0], 1.0 + 0.15 * (1.625 + 0.625))
    assert close(out["velocity"][0][1], 0.15 * 0.625)
    assert close(out["velocity"][1][0], 0.15 * (-1.625 + 0.625))
    assert close(out["velocity"][1][1], 1.0 + 0.15 * 0.625)


def test_boids_interaction_radius_inclusive():
    # at d = 10 exactly the other boid counts (cohesion 5 * 0.65 = 3.25);
    # at d = 10.001 it does not (the boid is alone: zero change)
    out = one_step("boids2d", BOIDS, [{"pos": (100.0, 100.0), "velocity": (0.0, 0.0)},
                                      {"pos": (110.0, 100.0), "velocity": (0.0, 0.0)},
                                      {"pos": (300.0, 300.0), "velocity": (0.0, 0.0)},
                                      {"pos": (310.001, 300.0), "velocity": (0.0, 0.0)}])
    assert close(out["velocity"][0][0], 0.4875)
    assert out["velocity"][2][0] == 0.0 and out["velocity"][3][0] == 0.0


def test_boids_bound_position_wraps():
    # |v| clamps to 2, x moves by 2 * 0.0005 = 0.001 past the bound and
    # boundPosition wraps it to the opposite bound
    out = one_step("boids2d", BOIDS, [{"pos": (999.9999, 500.0), "velocity": (2.0, 0.0)},
                                      {"pos": (0.0001, 200.0), "velocity": (-2.0, 0.0)}])
    assert out["pos"][0][0] == 0.0 and out["pos"][1][0] == 1000.0
    assert close(out["velocity"][0][0], 2.0) and close(out["velocity"][1][0], -2.0)


# ---- predator_prey (examples/predator_prey.abl:99-309); steps indexed in
# the simulate statement's order ------------------------------------------
PRED, PREY, GRASS = 0, 1, 2
S_PREY_MOVE, S_PRED_MOVE, S_PREY_EAT, S_PRED_EAT, S_GRASS_EATEN, S_PREY_EATEN = 4, 5, 6, 7, 8, 9
S_PRED_REPRO, S_PREY_REPRO, S_GRASS_GROWTH = 10, 11, 12


def pp(**kw):
    return predator_prey_params(num_agents=1600, agent_density=1600.0, **kw)  # env 1 x 1


def pp_state(o, pred=(), prey=(), grass=()):
    for t, recs in ((PRED, pred), (PREY, prey), (GRASS, grass)):
        arr = np.zeros(len(recs), dtype=o.dtypes[t])
        for i, r in enumerate(recs):
            for k, v in r.items():
                arr[i][k] = v
        o.set_agents(arr, t)


def test_pp_prey_eaten_strictly_inside_kill_distance():
    # predator at (0.5, 0.5); prey at 0.01 (< 0.02) is removed, prey at 0.03 stays
    o = Oracle("predator_prey").set(**pp())
    pp_state(o, pred=[{"pos": (0.5, 0.5), "life": 50}],
             prey=[{"pos": (0.51, 0.5), "life": 50}, {"pos": (0.5, 0.53), "life": 50}])
    o.run_step(S_PREY_EATEN)
    left = o.agents(PREY)
    assert len(left) == 1 and tuple(left["pos"][0]) == (0.5, 0.53)


def test_pp_predator_eats_every_prey_in_reach_and_starves():
    # life 10 + 2 prey within 0.02 * 35 = 80; a predator at life 0 with no
    # prey in reach is removed (life < 1)
    o = Oracle("predator_prey").set(**pp())
    pp_state(o, pred=[{"pos": (0.2, 0.2), "life": 10}, {"pos": (0.8, 0.8), "life": 0}],
             prey=[{"pos": (0.21, 0.2), "life": 5}, {"pos": (0.2, 0.19), "life": 5},
                   {"pos": (0.4, 0.4), "life": 5}])
    o.run_step(S_PRED_EAT)
    pred = o.agents(PRED)
    assert len(pred) == 1 and pred["life"][0] == 80
    assert len(o.agents(PREY)) == 3  # prey are removed by prey_eaten, not here


def test_pp_prey_eats_available_grass_and_starves():
    # life 5 + 2 available grass in reach * 60 = 125 (the unavailable one
    # gives nothing); a prey at life 0 with no grass is removed
    o = Oracle("predator_prey").set(**pp())
    pp_state(o, prey=[{"pos": (0.3, 0.3), "life": 5}, {"pos": (0.9, 0.1), "life": 0}],
             grass=[{"pos": (0.31, 0.3), "avail": True}, {"pos": (0.3, 0.29), "avail": True},
                    {"pos": (0.29, 0.3), "avail": False}])
    o.run_step(S_PREY_EAT)
    prey = o.agents(PREY)
    assert len(prey) == 1 and prey["life"][0] == 125


def test_pp_grass_eaten_only_when_available():
    o = Oracle("predator_prey").set(**pp())
    pp_state(o, prey=[{"pos": (0.5, 0.5), "life": 5}],
             grass=[{"pos": (0.51, 0.5), "avail": True, "dead_cycles": 7},
                    {"pos": (0.5, 0.51), "avail": False, "dead_cycles": 7},
                    {"pos": (0.9, 0.9), "avail": True, "dead_cycles": 7}])
    o.run_step(S_GRASS_EATEN)
    g = o.agents(GRASS)
    assert list(g["avail"]) == [False, False, True]
    assert list(g["dead_cycles"]) == [0, 7, 7]


def test_pp_grass_growth_two_ifs():
    # dead at GRASS_REGROW_CYCLES: the first if revives it (dead_cycles 0),
    # the second (in.avail is false) then sets dead_cycles = 60 + 1
    o = Oracle("predator_prey").set(**pp())
    pp_state(o, grass=[{"pos": (0.1, 0.1), "avail": False, "dead_cycles": 60},
                       {"pos": (0.2, 0.1), "avail": False, "dead_cycles": 10},
                       {"pos": (0.3, 0.1), "avail": True, "dead_cycles": 3}])
    o.run_step(S_GRASS_GROWTH)
    g = o.agents(GRASS)
    assert list(g["avail"]) == [True, False, True]
    assert list(g["dead_cycles"]) == [61, 11, 3]


@pytest.mark.parametrize("prob,births", [(1.0, 2), (0.0, 0)])
def test_pp_predator_reproduction_halves_life(prob, births):
    # with probability 1 every predator spawns: parent and child each keep
    # life / 2 (integer division: 9 -> 4, 8 -> 4), the child at the parent's
    # position with the opposite direction and steer
    o = Oracle("predator_prey").set(**pp(REPRODUCE_PREDATOR_PROB=prob))
    pp_state(o, pred=[{"pos": (0.2, 0.3), "dir": (0.3, 0.4), "steer": (0.1, 0.2), "life": 9},
                      {"pos": (0.7, 0.6), "dir": (-0.6, 0.8), "steer": (0.0, -0.5), "life": 8}])
    o.run_step(S_PRED_REPRO)
    p = o.agents(PRED)
    assert len(p) == 2 + births
    if births:
        assert list(p["life"]) == [4, 4, 4, 4]
        for parent, child in ((0, 2), (1, 3)):
            assert tuple(p["pos"][child]) == tuple(p["pos"][parent])
            assert tuple(p["dir"][child]) == tuple(-p["dir"][parent])
            assert tuple(p["steer"][child]) == tuple(-p["steer"][parent])
    else:
        assert list(p["life"]) == [9, 8]


def test_pp_prey_move_speed_limit_and_life():
    # dir + steer = (2, 0) is longer than 1: normalised to (1, 0), so x moves
    # by DELTA_TIME = 0.0015; a unit-speed prey keeps its speed; life - 1
    o = Oracle("predator_prey").set(**pp())
    pp_state(o, prey=[{"pos": (0.5, 0.5), "dir": (1.0, 0.0), "steer": (1.0, 0.0), "life": 7},
                      {"pos": (0.2, 0.2), "dir": (0.6, 0.8), "steer": (0.0, 0.0), "life": 3}])
    o.run_step(S_PREY_MOVE)
    p = o.agents(PREY)
    assert close(p["pos"][0][0], 0.5015) and p["pos"][0][1] == 0.5
    assert close(p["dir"][0][0], 1.0)
    assert close(p["pos"][1][0], 0.2 + 0.6 * 0.0015) and close(p["pos"][1][1], 0.2 + 0.8 * 0.0015)
    assert list(p["life"]) == [6, 2]


def test_pp_predator_moves_twice_as_fast():
    o = Oracle("predator_prey").set(**pp())
    pp_state(o, pred=[{"pos": (0.5, 0.5), "dir": (0.0, 1.0), "steer": (0.0, 0.0), "life": 7}])
    o.run_step(S_PRED_MOVE)
    p = o.agents(PRED)
    assert close(p["pos"][0][1], 0.503) and p["life"][0] == 6


def test_pp_bound_position_clamps_to_env():
    # boundPosition is clamp(pos, env_size): a prey leaving the 1 x 1
    # environment stays on its edge
    o = Oracle("predator_prey").set(**pp())
    pp_state(o, prey=[{"pos": (0.9995, 0.0005), "dir": (0.6, -0.8), "steer": (0.0, 0.0),
                       "life": 3}])
    o.run_step(S_PREY_MOVE)
    p = o.agents(PREY)
    assert tuple(p["pos"][0]) == (1.0, 0.0)
