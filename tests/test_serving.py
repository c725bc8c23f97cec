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
    from openabl_amd.serve import Server, estimate_footprint, sizing
    params = {"num_agents": 200000, "agent_density": 0.065536}
    n, cells = sizing("boids2d", params)
    one = estimate_footprint("boids2d", "f64", params, n, cells)
    budget = int(2.5 * one)
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


FOOTPRINT_CASES = [
    ("boids2d", "f64", {"num_agents": 1 << 20, "agent_density": 0.065536}, 30),
    ("boids2d", "f32", {"num_agents": 65536, "agent_density": 0.065536}, 30),
    ("circle", "f64", {"num_agents": 100000, "rho": 0.1024, "r": 2.5}, 10),
    ("circle3d", "f64", {"num_agents": 262144, "rho": 262144 / 1e9, "r": 4.0,
                         "num_clusters": 16, "cluster_sigma": 20.0}, 3),
    ("game_of_life", "f64", {"num_agents": 1024 * 1024, "alive_fraction": 0.3}, 5),
    ("predator_prey", "f64", {"num_agents": 100000, "prey_fraction": 0.4,
                              "predator_fraction": 0.04}, 20),
]


@pytest.mark.parametrize("prog,prec,params,steps", FOOTPRINT_CASES,
                         ids=[f"{c[0]}-{c[1]}" for c in FOOTPRINT_CASES])
def test_estimate_bounds_live_footprint(prog, prec, params, steps):
    """The admission estimate is an upper bound of what a live simulation
    holds after its grid, neighbour lists and (for predator_prey) flag and
    newborn buffers exist: abl_hip_footprint <= abl_hip_estimate_footprint."""
    from openabl_amd import Simulation, load_library
    from openabl_amd.serve import estimate_footprint, sizing
    print(f"case footprint {prog} {prec}", flush=True)
    load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    sim = Simulation(prog, prec, params=params)
    sim.init_agents()
    sim.simulate(steps)
    sim.synchronize()
    live = sim.footprint()
    sim.close()
    n, cells = sizing(prog, params)
    est = estimate_footprint(prog, prec, params, n, cells)
    print(f"  live {live} estimate {est} ({est / live:.2f}x)", flush=True)
    assert live <= est


def test_status_polling_while_runs_finish():
    """One thread polls status, run times and stats while runs start and
    finish on the server's threads: every run seen DONE already has its end
    time and results (latencies non-negative, admit <= done), and the final
    statistics agree with the runs' own times."""
    import threading
    from openabl_amd.serve import Server
    srv = Server(hbm_budget=8 << 30, max_active=3, chunk_steps=5)
    rids = [srv.submit("circle", "f64", {"num_agents": 2000 + 500 * k, "rho": 0.1024, "r": 2.5},
                       20 + 10 * k, (0xDEADBEEF + k, 0xBEEFDEAD)) for k in range(8)]
    seen, bad = set(), []
    stop = threading.Event()

    def poll():
        while not stop.is_set():
            for rid in rids:
                if srv.status(rid) == 2:
                    t = srv.times(rid)
                    if not (t[0] >= 0 and t[1] >= 0 and abs(t[2] - t[0] - t[1]) < 1e-6):
                        bad.append((rid, t))
                    if len(srv.result(rid)) == 0:
                        bad.append((rid, "no result"))
                    seen.add(rid)
            st = srv.stats()
            if st["completed"] and not (st["p50_s"] > 0 and st["p99_s"] >= st["p50_s"]):
                bad.append(("stats", st))

    th = threading.Thread(target=poll)
    th.start()
    for rid in rids:
        srv.wait(rid)
    stop.set()
    th.join()
    assert not bad, bad[:3]
    st = srv.stats()
    lat = sorted(srv.times(r)[2] for r in rids)
    assert st["completed"] == len(rids)
    assert abs(st["p50_s"] - lat[3]) < 1e-9 and abs(st["p99_s"] - lat[-1]) < 1e-9
    srv.close()
