"""Batched ensembles: every member of a batch is bit-identical to the same
run computed alone (same parameters, generator start state, cell scale, tile
shape and list settings), with and without graph replay, with neighbour lists
(members re-sorted and rebuilt on their own schedules) and without, FP64 and
FP32."""
import math

import numpy as np
import pytest

from orclib import field_bytes, fold6

pytestmark = pytest.mark.gpu


def members(k, seed=5):
    rng = np.random.default_rng(seed)
    out = []
    for n in np.exp(rng.uniform(math.log(1000), math.log(8000), k)).astype(int):
        w = rng.uniform(0.5, 2.0, 3)
        out.append({"num_agents": int(n), "agent_density": 0.065,
                    "max_pos": fold6(math.sqrt(n / 0.065)),
                    "steer_scale": float(w[0]), "collision_scale": float(w[1]),
                    "match_scale": float(w[2])})
    return out


LISTS = {"stencil": {"hip.nlist": 0},
         # a fixed list capacity, so a list that outgrows it overflows in the
         # ensemble exactly when it does in the run alone
         "lists": {"hip.nlist": 1, "hip.list_cap": 96},
         # a skin a boid at full speed crosses in three steps: rebuilds
         "lists-skin": {"hip.nlist": 1, "hip.list_cap": 96, "hip.skin": 0.0005}}


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("graph,steps", [(1, 30), (1, 31), (0, 12)])
@pytest.mark.parametrize("mode", sorted(LISTS))
def test_members_equal_alone_runs(prec, graph, steps, mode):
    from openabl_amd import MEMBERS, Simulation, load_library, record_dtype
    from openabl_amd.ensemble import Ensemble, member_seed
    print(f"case ensemble members vs alone {prec} graph={graph} steps={steps} {mode}", flush=True)
    ms = members(10)
    cfg = dict({"hip.cell_scale": 2, "hip.tile": 1}, **LISTS[mode])
    ens = Ensemble("boids2d", prec, ms, config=dict(cfg, **{"hip.graph": graph}))
    ens.init(0x1234, 0x5678)
    ens.simulate(steps)
    ens.synchronize()
    dt = record_dtype(MEMBERS["boids2d"], prec)
    for m, p in enumerate(ms):
        load_library().abl_rng_seed(*member_seed(0x1234, 0x5678, m))
        sim = Simulation("boids2d", prec, params=p, config=cfg)
        sim.init_agents()
        sim.simulate(steps)
        assert field_bytes(ens.download(m)) == field_bytes(sim.download(dt)), m
        sim.close()
    ens.close()


def test_lists_follow_each_members_own_rebuilds():
    """Members whose agents move at different speeds rebuild on different
    timesteps: half the members move 50x slower (time_scale), so with a skin
    the fast ones are due every step and the slow ones every few steps; every
    member still equals its run alone, which rebuilds only when its own agents
    stray."""
    from openabl_amd import MEMBERS, Simulation, load_library, record_dtype
    from openabl_amd.ensemble import Ensemble, member_seed
    print("case ensemble lists with members on different rebuild schedules", flush=True)
    ms = members(8, seed=9)
    for k, m in enumerate(ms):
        m["time_scale"] = 0.25 if k % 2 else 0.005
    cfg = {"hip.cell_scale": 2, "hip.tile": 1, "hip.nlist": 1, "hip.list_cap": 96,
           "hip.skin": 0.02}
    ens = Ensemble("boids2d", "f64", ms, config=cfg)
    ens.init(0x1234, 0x5678)
    ens.simulate(24)
    ens.synchronize()
    dt = record_dtype(MEMBERS["boids2d"], "f64")
    for m, p in enumerate(ms):
        load_library().abl_rng_seed(*member_seed(0x1234, 0x5678, m))
        sim = Simulation("boids2d", "f64", params=p, config=cfg)
        sim.init_agents()
        sim.simulate(24)
        assert field_bytes(ens.download(m)) == field_bytes(sim.download(dt)), m
        sim.close()
    ens.close()
