"""Batched ensembles: every member of a batch is bit-identical to the same
run computed alone (same parameters, generator start state, cell scale and
tile shape), with and without graph replay, FP64 and FP32."""
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


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("graph,steps", [(1, 30), (1, 31), (0, 12)])
def test_members_equal_alone_runs(prec, graph, steps):
    from openabl_amd import MEMBERS, Simulation, load_library, record_dtype
    from openabl_amd.ensemble import Ensemble, member_seed
    ms = members(10)
    cfg = {"hip.cell_scale": 2, "hip.tile": 1}
    ens = Ensemble("boids2d", prec, ms, config=dict(cfg, **{"hip.graph": graph}))
    ens.init(0x1234, 0x5678)
    ens.simulate(steps)
    ens.synchronize()
    dt = record_dtype(MEMBERS["boids2d"], prec)
    for m, p in enumerate(ms):
        load_library().abl_rng_seed(*member_seed(0x1234, 0x5678, m))
        # the ensemble walks stencils; so does this run (no neighbour lists)
        sim = Simulation("boids2d", prec, params=p, config=dict(cfg, **{"hip.nlist": 0}))
        sim.init_agents()
        sim.simulate(steps)
        assert field_bytes(ens.download(m)) == field_bytes(sim.download(dt)), m
        sim.close()
    ens.close()
