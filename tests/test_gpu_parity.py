"""GPU parity tests: the HIP path through the C ABI against the oracle.

Bars (stated per test):
  * initial states built by the -b hip host runtime: bit-exact;
  * neighbour sets (which agents each near loop accepts): bit-exact, from the
    same input state, at every size tested (up to 1M agents);
  * integer models (game_of_life): bit-exact over 100 steps;
  * FP64 state after one step from a shared state: |gpu - oracle| <=
    1e-12 * max(|oracle|, 1) per component when neighbours are visited in
    cell order (the default); bit-exact with hip.order=index, which visits
    them in the reference's index order;
  * over many steps: bit-exact in index order; in cell order the stated bound
    of each test (chaotic models drift by summation order only).
"""
import numpy as np
import pytest

from orclib import (Oracle, boids2d_params, circle3d_params, circle_params, field_bytes,
                    game_of_life_params)

pytestmark = pytest.mark.gpu


def make(model, prec, params, order="cell", init=True):
    from openabl_amd import Simulation
    sim = Simulation(model, prec, params=params, config={"hip.order": order})
    if init:
        sim.init_agents()
    return sim


def dtype_of(model, prec):
    from openabl_amd import MEMBERS, record_dtype
    return record_dtype(MEMBERS[model], prec)


def state(sim, model, prec):
    return sim.download(dtype_of(model, prec))


def assert_close(got, want, rel=1e-12):
    for name in want.dtype.names:
        a = got[name].astype(np.float64)
        b = want[name].astype(np.float64)
        tol = rel * np.maximum(np.abs(b), 1.0)
        bad = np.abs(a - b) > tol
        assert not bad.any(), (name, int(bad.sum()), float(np.max(np.abs(a - b))))


CASES = [
    ("circle", circle_params(num_agents=1024)),
    ("boids2d", boids2d_params(num_agents=4096)),
    ("circle3d", circle3d_params(num_agents=4000)),
    ("circle3d", circle3d_params(num_agents=8192, rho=8192 / 1e6, r=4.0, num_clusters=8,
                                 cluster_sigma=20.0)),
    ("game_of_life", game_of_life_params(num_agents=64 * 64, alive_fraction=0.3)),
]
IDS = ["circle", "boids2d", "circle3d", "circle3d-clustered", "game_of_life"]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("model,params", CASES, ids=IDS)
def test_initial_state_bit_exact(model, params, prec):
    """Host init of -b hip (abl_hip_init_agents) == the C backend's main()."""
    print(f"case init {model} {prec} n={params['num_agents']}", flush=True)
    from openabl_amd import load_library
    load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    sim = make(model, prec, params)
    o = Oracle(model, prec).set(**params).init()
    assert field_bytes(state(sim, model, prec)) == field_bytes(o.agents())


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("model,params", CASES, ids=IDS)
def test_neighbour_sets_bit_exact(model, params, prec):
    """Neighbour sets from the same state, initially and after 10 GPU steps."""
    print(f"case nbrs {model} {prec} n={params['num_agents']}", flush=True)
    from openabl_amd import load_library
    load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    sim = make(model, prec, params)
    o = Oracle(model, prec).set(**params).init()
    for _ in range(2):
        c, n = sim.neighbours(0)
        oc, on = o.neighbours(0)
        assert len(c) == len(oc)
        assert np.array_equal(c, oc) and np.array_equal(n, on)
        sim.simulate(10)
        o.set_agents(state(sim, model, prec))


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("model,params", CASES, ids=IDS)
def test_one_step_from_shared_state(model, params, prec):
    """One step from the same state: index order bit-exact; cell order within
    1e-12 relative (FP64) / 1e-4 relative (FP32).  The FP32 bound follows from
    reordering a float sum of k ~ 20 neighbour positions of magnitude <= 1000:
    |error of the mean| <= k * 2^-24 * 1000 ~ 1.2e-3, times global_scale 0.15."""
    print(f"case step {model} {prec}", flush=True)
    from openabl_amd import load_library
    load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    o = Oracle(model, prec).set(**params).init()
    o.simulate(3)
    s0 = o.agents()
    o.simulate(1)
    want = o.agents()
    for order in ("index", "cell"):
        sim = make(model, prec, params, order=order, init=False)
        sim.upload(s0)
        sim.simulate(1)
        got = state(sim, model, prec)
        if order == "index" or model == "game_of_life":
            assert field_bytes(got) == field_bytes(want), order
        else:
            assert_close(got, want, 1e-12 if prec == "f64" else 1e-4)


@pytest.mark.parametrize("model,params,steps", [
    ("circle", circle_params(num_agents=1024), 100),
    ("boids2d", boids2d_params(num_agents=2048), 50),
    ("circle3d", circle3d_params(num_agents=2000), 20),
    ("game_of_life", game_of_life_params(num_agents=64 * 64, alive_fraction=0.3), 100),
], ids=["circle", "boids2d", "circle3d", "game_of_life"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_long_run_index_order_bit_exact(model, params, steps, prec, tmp_path):
    """hip.order=index reproduces the C backend's run bit for bit: the final
    save() text is identical (here for the configured step counts)."""
    print(f"case long {model} {prec} {steps}", flush=True)
    from openabl_amd import load_library
    load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    sim = make(model, prec, params, order="index")
    o = Oracle(model, prec).set(**params).init()
    sim.simulate(steps)
    o.simulate(steps)
    assert field_bytes(state(sim, model, prec)) == field_bytes(o.agents())


def test_circle_cell_order_100_steps_bound():
    """Cell order over the full circle config (1,024 agents, 100 steps, FP64):
    summation order is the only difference (index order is bit-exact, test
    above).  Per-step differences of ~1 ulp are amplified by the dynamics
    (about x1.25 per step, measured: 8e-6 after 100 steps), so the stated
    bound is 1e-4 absolute on positions in the 100-wide box (1e-6 relative),
    and <= 1e-12 after the first step."""
    params = circle_params(num_agents=1024)
    from openabl_amd import load_library
    load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    sim = make("circle", "f64", params)
    o = Oracle("circle").set(**params).init()
    done = 0
    for upto in (1, 10, 100):
        sim.simulate(upto - done)
        o.simulate(upto - done)
        done = upto
        d = np.abs(state(sim, "circle", "f64")["pos"] - o.agents()["pos"]).max()
        print(f"circle cell order: max |dpos| after {upto} steps = {d:.3e}")
        if upto == 1:
            assert d <= 1e-12 * 100
    assert d <= 1e-4


def test_game_of_life_1m_bit_exact_vs_conway():
    """1024 x 1024 game_of_life (alive 0.3), 20 steps: bit-exact against the
    Conway stencil (which the oracle matches, test_oracle.py)."""
    from test_oracle import conway
    from openabl_amd import load_library
    load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    params = game_of_life_params(num_agents=1024 * 1024, alive_fraction=0.3)
    sim = make("game_of_life", "f64", params)
    grid = state(sim, "game_of_life", "f64")["alive"].astype(bool).reshape(1024, 1024)
    sim.simulate(20)
    for _ in range(20):
        grid = conway(grid)
    got = state(sim, "game_of_life", "f64")["alive"].astype(bool).reshape(1024, 1024)
    assert np.array_equal(got, grid)


def test_boids_1m_neighbour_sets_bit_exact():
    """The benchmark workload (1,048,576 boids in a 4000 x 4000 box, radius 10):
    neighbour sets bit-exact against the oracle's grid path."""
    params = boids2d_params(num_agents=1 << 20)
    from openabl_amd import load_library
    load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    sim = make("boids2d", "f64", params)
    sim.simulate(3)
    s = state(sim, "boids2d", "f64")
    o = Oracle("boids2d").set(**params)
    o.set_agents(s)
    c, n = sim.neighbours(0)
    oc, on = o.neighbours(0)
    assert np.array_equal(c, oc) and np.array_equal(n, on)


@pytest.mark.parametrize("model", ["circle", "boids2d"])
def test_edge_cases(model):
    """Empty population, one agent, and agents stacked on one point / on the
    environment's max corner (cell index clamp)."""
    from openabl_amd import Simulation
    dt = dtype_of(model, "f64")
    params = circle_params(num_agents=64) if model == "circle" else boids2d_params(num_agents=64)
    # empty
    sim = Simulation(model, "f64", params=params)
    sim.upload(np.zeros(0, dtype=dt))
    sim.simulate(3)
    assert sim.count() == 0 and len(sim.neighbours(0)[0]) == 0
    # stacked + corner + single
    hi = params["W"] if model == "circle" else params["max_pos"]
    for pts in ([(5.0, 5.0)] * 7 + [(hi, hi)] * 3 + [(0.0, hi)], [(hi / 2, hi / 3)]):
        recs = np.zeros(len(pts), dtype=dt)
        recs["pos"] = np.array(pts)
        if model == "boids2d":
            recs["velocity"] = np.linspace(-1, 1, 2 * len(pts)).reshape(-1, 2)
        o = Oracle(model).set(**params)
        o.set_agents(recs)
        for order in ("index", "cell"):
            sim = Simulation(model, "f64", params=params, config={"hip.order": order})
            sim.upload(recs)
            c, n = sim.neighbours(0)
            oc, on = o.neighbours(0)
            assert np.array_equal(c, oc) and np.array_equal(n, on)
            sim.simulate(1)
            o2 = Oracle(model).set(**params)
            o2.set_agents(recs)
            o2.simulate(1)
            assert field_bytes(state(sim, model, "f64")) == field_bytes(o2.agents())


def test_errors_are_reported():
    from openabl_amd import AblError, Simulation
    with pytest.raises(AblError):
        Simulation("no_such_model")
    sim = Simulation("circle", "f64")
    with pytest.raises(AblError):
        sim.set_param("nonexistent", 1.0)
    with pytest.raises(AblError):
        sim.set_config("hip.block", "100")


def _float_threshold_le(r: float) -> np.float32:
    """max float32 f >= 0 with float(sqrt(f)) <= r (numpy's float32 sqrt is
    correctly rounded)."""
    lo, hi = 0, 0x7F800000
    while hi - lo > 1:
        mid = (lo + hi) // 2
        f = np.array([mid], dtype=np.uint32).view(np.float32)[0]
        if float(np.sqrt(f)) <= r:
            lo = mid
        else:
            hi = mid
    return np.array([lo], dtype=np.uint32).view(np.float32)[0]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("order", ["cell", "index"])
def test_boids_radius_boundary(prec, order):
    """Isolated pairs of boids whose squared distance straddles the float
    rounding boundary of the interaction radius (10) and of the separation
    radius (2): the update_boid kernel accepts exactly the pairs the C
    backend's (double) sqrtf((float) s) <= r (resp. < r) accepts, so one step
    is bit-exact (two-term sums do not depend on order)."""
    params = boids2d_params(num_agents=65536)  # 1000 x 1000 box
    dt = dtype_of("boids2d", prec)
    ft = np.float64 if prec == "f64" else np.float32
    offsets = []
    for r in (10.0, 2.0):
        t = _float_threshold_le(r)
        t_up = np.nextafter(t, np.float32(np.inf))
        for target in (float(t), (float(t) + float(t_up)) / 2, float(t_up)):
            d = ft(np.sqrt(target))
            for j in range(-3, 4):
                dd = d
                for _ in range(abs(j)):
                    dd = np.nextafter(dd, ft(np.inf) if j > 0 else ft(-np.inf))
                offsets.append(dd)
    pts, vel = [], []
    rng = np.random.default_rng(11)
    # A on the x = 0 edge so that B's x coordinate is the offset itself
    # (no rounding of x0 + dx), pairs 20 apart in y
    for k, dx in enumerate(offsets):
        y0 = ft(10 + 20 * k)
        pts += [(ft(0), y0), (dx, y0)]
        vel += [tuple(rng.uniform(-1, 1, 2)), tuple(rng.uniform(-1, 1, 2))]
    assert 10 + 20 * len(offsets) <= params["max_pos"]
    recs = np.zeros(len(pts), dtype=dt)
    recs["pos"] = np.array(pts, dtype=ft)
    recs["velocity"] = np.array(vel, dtype=ft)
    o = Oracle("boids2d", prec).set(**params)
    o.set_agents(recs)
    o.simulate(1)
    from openabl_amd import Simulation
    sim = Simulation("boids2d", prec, params=params, config={"hip.order": order})
    sim.upload(recs)
    sim.simulate(1)
    assert field_bytes(state(sim, "boids2d", prec)) == field_bytes(o.agents())
