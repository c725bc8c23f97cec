"""Serving: continuous batching of runs of different models on one GPU.

Each run's result must be bit-identical to the same run computed alone (same
parameters, generator state and configuration), whatever else was admitted
beside it; admission must respect the device-memory budget.
"""
import numpy as np
import pytest

from orclib import field_bytes

pytestmark = pytest.mark.gpu

MIX = [
    ("circle", "f64", {"num_agents": 3000, "rho": 0.1024, "r": 2.5}, 120),
    ("boids2d", "f32", {"num_agents": 20000, "agent_density": 0.065536}, 150),
    ("game_of_life", "f64", {"num_agents": 128 * 128, "alive_fraction": 0.3}, 200),
    ("predator_prey", "f64", {"num_agents": 8000, "prey_fraction": 0.4,
                              "predator_fraction": 0.04}, 100),
    ("circle", "f64", {"num_agents": 1500, "rho": 0.05, "r": 5.0}, 300),
    ("boids2d", "f64", {"num_agents": 5000, "agent_density": 0.065536}, 90),
]


def alone(program, precision, params, steps, rng):
    from openabl_amd import AGENTS, Simulation, load_library, record_dtype
    load_library().abl_rng_seed(*rng)
    sim = Simulation(program, precision, params=params)
    sim.init_agents()
    sim.simulate(steps)
    out = [sim.download(record_dtype(m, precision), t) for t, (_, m) in enumerate(AGENTS[program])]
    log = sim.log_text()
    sim.close()
    return out, log


def test_mix_results_equal_alone_runs():
    from openabl_amd import AGENTS
    from openabl_amd.serve import Server
    srv = Server(hbm_budget=6 << 30, max_active=4, chunk_steps=25)
    ids = []
    for k, (prog, prec, params, steps) in enumerate(MIX):
        rng = (0xDEADBEEF + k, 0xBEEFDEAD + 7 * k)
        ids.append((srv.submit(prog, prec, params, steps, rng), rng))
    for rid, _ in ids:
        srv.wait(rid)
    stats = srv.stats()
    print("serving stats", stats, flush=True)
    assert stats["completed"] == len(MIX)
    assert stats["p99_s"] >= stats["p50_s"] > 0
    for (prog, prec, params, steps), (rid, rng) in zip(MIX, ids):
        want, want_log = alone(prog, prec, params, steps, rng)
        for t in range(len(AGENTS[prog])):
            assert field_bytes(srv.result(rid, t)) == field_bytes(want[t]), (prog, t)
        if prog == "predator_prey":
            got = [r.split(",")[:3] for r in srv.log(rid).splitlines()]
            assert got == [r.split(",")[:3] for r in want_log.splitlines()]
    srv.close()


def test_admission_respects_budget():
    """A budget that holds about two of these runs at once: all complete,
    the peak admitted footprint never exceeds the budget."""
    from openabl_amd.serve import Server, sizing
    from openabl_amd import load_library
    n, cells = sizing("boids2d", {"num_agents": 200000, "agent_density": 0.065536})
    import ctypes as C
    one = C.c_size_t()
    load_library().abl_hip_estimate_footprint(b"boids2d", 0, n, cells, C.byref(one))
    budget = int(2.5 * one.value)
    srv = Server(hbm_budget=budget, max_active=8, chunk_steps=10)
    rids = [srv.submit("boids2d", "f64", {"num_agents": 200000, "agent_density": 0.065536}, 40)
            for _ in range(5)]
    for r in rids:
        srv.wait(r)
    st = srv.stats()
    assert st["completed"] == 5
    assert st["peak_bytes"] <= budget
    srv.close()


def test_oversized_run_fails_cleanly():
    from openabl_amd import AblError
    from openabl_amd.serve import Server
    srv = Server(hbm_budget=1 << 20, max_active=2, chunk_steps=10)
    rid = srv.submit("boids2d", "f64", {"num_agents": 100000, "agent_density": 0.065536}, 10)
    with pytest.raises(AblError):
        srv.wait(rid)
    assert srv.status(rid) == 3
    srv.close()
