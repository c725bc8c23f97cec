"""predator_prey: run-time add()/removeCurrent(), random() inside steps,
cross-type near loops and the sequential step (count, sum, log_csv).

The reference C backend rejects this model (src/backend/CBackend.cpp:30-31),
so no reference run exists.  The oracle (oracle/models/predator_prey.c with
oracle/orc_dynamic.c) states the -b hip semantics; it is pinned here by
invariants (grid path == brute force, population bookkeeping, Conway-free
closed checks below), and the GPU path must reproduce it bit for bit when
neighbours are visited in the reference's index order.
"""
import numpy as np
import pytest

from orclib import Oracle, field_bytes, predator_prey_params

TYPES = ["Predator", "Prey", "Grass"]


def parse_log(text):
    rows = [r.split(",") for r in text.strip().splitlines()]
    return [(int(a), int(b), int(c)) for a, b, c, _ in rows]


def test_oracle_grid_equals_brute_force():
    p = predator_prey_params(num_agents=1600, agent_density=400.0)
    a = Oracle("predator_prey", grid=False).set(**p).init()
    b = Oracle("predator_prey", grid=True).set(**p).init()
    a.simulate(12)
    b.simulate(12)
    for t in range(3):
        assert field_bytes(a.agents(t)) == field_bytes(b.agents(t)), TYPES[t]
        assert np.array_equal(a.ids(t), b.ids(t))
    assert a.log_text() == b.log_text()


def test_oracle_population_bookkeeping():
    """Ids stay unique and ascending; newborn ids continue each type's
    sequence; the logged counts are the live populations; grass is never
    added or removed; every predator/prey life stays >= 1 after the
    starve steps (removeCurrent() on life < 1)."""
    p = predator_prey_params(num_agents=6400, agent_density=1600.0,
                             REPRODUCE_PREY_PROB=0.05, REPRODUCE_PREDATOR_PROB=0.05)
    o = Oracle("predator_prey").set(**p).init()
    n0 = [len(o.agents(t)) for t in range(3)]
    births = 0
    for step in range(15):
        o.simulate(1)
        for t in range(3):
            ids = o.ids(t)
            assert np.all(np.diff(ids) > 0)
        births += int((o.ids(1) >= n0[1]).sum())
        assert len(o.agents(2)) == n0[2]
        counts = parse_log(o.log_text())[-1]
        assert counts[0] == len(o.agents(1)) and counts[1] == len(o.agents(0))
        assert counts[2] == int(o.agents(2)["avail"].sum())
    assert births > 0
    assert len(o.agents(1)) != n0[1]


# ---- GPU -------------------------------------------------------------------------


def gpu_sim(params, order="index"):
    from openabl_amd import Simulation, load_library
    load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    sim = Simulation("predator_prey", "f64", params=params, config={"hip.order": order})
    sim.init_agents()
    return sim


def dtypes(prec="f64"):
    from openabl_amd import AGENTS, record_dtype
    return [record_dtype(m, prec) for _, m in AGENTS["predator_prey"]]


@pytest.mark.gpu
def test_gpu_initial_state_bit_exact():
    p = predator_prey_params(num_agents=3200)
    sim = gpu_sim(p)
    o = Oracle("predator_prey").set(**p).init()
    for t, dt in enumerate(dtypes()):
        assert field_bytes(sim.download(dt, t)) == field_bytes(o.agents(t)), TYPES[t]


@pytest.mark.gpu
@pytest.mark.parametrize("n,steps", [(3200, 40), (100000, 6)], ids=["3200x40", "100k-serving-x6"])
def test_gpu_index_order_bit_exact(n, steps):
    """Agents, ids and the log's integer columns after `steps` timesteps."""
    print(f"case predator_prey n={n} steps={steps}", flush=True)
    if n == 100000:
        p = predator_prey_params(num_agents=n, prey_fraction=0.4, predator_fraction=0.04)
    else:
        p = predator_prey_params(num_agents=n, REPRODUCE_PREY_PROB=0.02,
                                 REPRODUCE_PREDATOR_PROB=0.02)
    sim = gpu_sim(p)
    o = Oracle("predator_prey").set(**p).init()
    for chunk in (1, steps - 1):
        sim.simulate(chunk)
        o.simulate(chunk)
        for t, dt in enumerate(dtypes()):
            got = sim.download(dt, t)
            assert np.array_equal(sim.ids(t), o.ids(t)), TYPES[t]
            assert field_bytes(got) == field_bytes(o.agents(t)), TYPES[t]
        assert parse_log(sim.log_text()) == parse_log(o.log_text())


@pytest.mark.gpu
def test_gpu_cell_order_deterministic_and_consistent():
    """Cell-ordered visiting: two runs agree bit for bit; the sequential
    step's counts equal the downloaded populations."""
    p = predator_prey_params(num_agents=20000, prey_fraction=0.4, predator_fraction=0.04)
    a = gpu_sim(p, "cell")
    b = gpu_sim(p, "cell")
    a.simulate(10)
    b.simulate(10)
    for t, dt in enumerate(dtypes()):
        assert field_bytes(a.download(dt, t)) == field_bytes(b.download(dt, t))
    rows = parse_log(a.log_text())
    assert len(rows) == 10
    assert rows[-1][0] == a.count(1) and rows[-1][1] == a.count(0)
    assert rows[-1][2] == int(a.download(dtypes()[2], 2)["avail"].sum())
    assert a.reduce_sum(2, "avail") == rows[-1][2]
    assert a.count_eq(2, "avail", 1) == rows[-1][2]
