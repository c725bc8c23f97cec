"""Neighbour lists (hip.nlist=1, the default for steps whose near loop runs
over their own type): per-agent lists of the halo records within radius +
skin, built right after a (re-)sort and reused while no agent has moved more
than skin / 2; rebuilds are decided on the device.

  * with a vanishing skin every timestep rebuilds, so every step walks lists
    built on that step's sorted order: runs are bit-identical to the stencil
    walk (hip.nlist=0);
  * lists too short to hold an agent's candidates (hip.list_cap) send it
    back to the stencil walk on a fresh grid: again bit-identical;
  * the gated re-sort and list build run as one launch whose phases a work
    queue orders (hip.rebuild=fused, the default) or as four launches
    (split): the two are bit-identical over runs that reuse and rebuild the
    lists, also when the launch has many times more blocks than the device
    holds at once and when several simulations rebuild at the same time;
  * with the default skin the lists are reused for many steps: the same
    neighbours are found (exact near test on every list entry), visited in
    the order of the last build, so results match the stencil walk to the
    summation-order level (over 20 boids steps in a 550-wide box: 1e-9
    absolute in FP64, 1e-6 with the float velocities of f32c) and the FP64
    oracle within the per-step bound of test_gpu_parity.
"""
import numpy as np
import pytest

from orclib import Oracle, boids2d_params, circle3d_params, circle_params, field_bytes

pytestmark = pytest.mark.gpu


def run(model, prec, params, steps, **cfg):
    from openabl_amd import MEMBERS, Simulation, load_library, record_dtype
    load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    sim = Simulation(model, prec, params=params, config=cfg)
    sim.init_agents()
    out = []
    done = 0
    for upto in steps:
        sim.simulate(upto - done)
        done = upto
        out.append(sim.download(record_dtype(MEMBERS[model], prec)))
    sim.close()
    return out


CASES = [
    ("boids2d", "f64", boids2d_params(num_agents=20000), {"hip.cell_scale": 2}),
    ("boids2d", "f32", boids2d_params(num_agents=20000), {"hip.cell_scale": 1}),
    ("boids2d", "f32c", boids2d_params(num_agents=20000), {"hip.cell_scale": 2, "hip.tile": 3}),
    ("circle", "f64", circle_params(num_agents=4096), {}),
    ("circle3d", "f64", circle3d_params(num_agents=32768, rho=32768 / 1e9, r=4.0,
                                        num_clusters=2, cluster_sigma=20.0), {}),
    ("circle3d", "f32c", circle3d_params(num_agents=8000), {"hip.cell_scale": 2}),
]
IDS = [f"{c[0]}-{c[1]}-{i}" for i, c in enumerate(CASES)]


@pytest.mark.parametrize("model,prec,params,cfg", CASES, ids=IDS)
def test_rebuild_every_step_equals_stencil(model, prec, params, cfg):
    print(f"case nlist tiny skin {model} {prec} {cfg}", flush=True)
    steps = (1, 2, 5, 12)
    # same cells for both (list mode would widen them by 1 + skin)
    a = run(model, prec, params, steps, **dict(cfg, **{"hip.skin": 1e-9, "hip.cell_inflate": 1}))
    b = run(model, prec, params, steps, **dict(cfg, **{"hip.nlist": 0}))
    for t, x, y in zip(steps, a, b):
        assert field_bytes(x) == field_bytes(y), t


@pytest.mark.parametrize("model,prec,params,cfg", CASES[:4], ids=IDS[:4])
def test_overflowing_lists_equal_stencil(model, prec, params, cfg):
    print(f"case nlist overflow {model} {prec} {cfg}", flush=True)
    steps = (1, 3, 6)
    a = run(model, prec, params, steps, **dict(cfg, **{"hip.list_cap": 4, "hip.cell_inflate": 1}))
    b = run(model, prec, params, steps, **dict(cfg, **{"hip.nlist": 0}))
    for t, x, y in zip(steps, a, b):
        assert field_bytes(x) == field_bytes(y), t


@pytest.mark.parametrize("prec", ["f64", "f32c"])
def test_reused_lists_match_stencil(prec):
    params = boids2d_params(num_agents=20000)
    steps = (1, 5, 20)
    a = run("boids2d", prec, params, steps, **{"hip.cell_scale": 2, "hip.cell_inflate": 1})
    b = run("boids2d", prec, params, steps, **{"hip.cell_scale": 2, "hip.nlist": 0})
    # the first step walks lists built on the same sorted order: identical
    assert field_bytes(a[0]) == field_bytes(b[0])
    for t, x, y in zip(steps, a, b):
        d = max(np.abs(x["pos"] - y["pos"]).max(), np.abs(x["velocity"] - y["velocity"]).max())
        print(f"boids2d {prec} lists vs stencil after {t} steps: {d:.3e}")
        # FP64: summation order only; f32c: float velocities (ulp 6e-8 at |v| ~ 1)
        assert d <= (1e-9 if prec == "f64" else 1e-6)


def test_lists_one_step_vs_oracle_after_reuse():
    """After 30 steps on reused lists, one more step from that state agrees
    with the FP64 oracle within 1e-12 relative."""
    from openabl_amd import MEMBERS, Simulation, load_library, record_dtype
    params = boids2d_params(num_agents=8192)
    dt = record_dtype(MEMBERS["boids2d"], "f64")
    load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    sim = Simulation("boids2d", "f64", params=params, config={"hip.cell_scale": 2})
    sim.init_agents()
    sim.simulate(30)
    s0 = sim.download(dt)
    sim.simulate(1)
    got = sim.download(dt)
    o = Oracle("boids2d").set(**params)
    o.set_agents(s0)
    o.simulate(1)
    want = o.agents()
    for f in ("pos", "velocity"):
        tol = 1e-12 * np.maximum(np.abs(want[f]), 1.0)
        assert (np.abs(got[f] - want[f]) <= tol).all(), f


@pytest.mark.parametrize("model,prec,params,cfg", CASES, ids=IDS)
def test_one_launch_rebuild_equals_split(model, prec, params, cfg):
    print(f"case nlist fused vs split rebuild {model} {prec} {cfg}", flush=True)
    steps = (1, 3, 8, 20)
    for skin in (1e-9, 0.05):
        a = run(model, prec, params, steps, **dict(cfg, **{"hip.skin": skin}))
        b = run(model, prec, params, steps, **dict(cfg, **{"hip.skin": skin, "hip.rebuild": "split"}))
        for t, x, y in zip(steps, a, b):
            assert field_bytes(x) == field_bytes(y), f"{model} {prec} skin {skin}: step {t} differs"


def test_one_launch_rebuild_oversubscribed_equals_split():
    """The one-launch rebuild's phase waits do not need its blocks resident
    together: with 65,528 blocks (far more than the CUs hold at once, so most
    start only after others have left) the run at the benchmark's launch
    configuration, over steps that reuse and rebuild the lists, is
    bit-identical to the four-launch rebuild."""
    print("case nlist oversubscribed one-launch rebuild", flush=True)
    params = boids2d_params(num_agents=65536)
    cfg = {"hip.cell_scale": 2, "hip.lds_slack": 1.25}
    steps = (1, 4, 15)
    for skin in (1e-9, 0.05):
        a = run("boids2d", "f64", params, steps,
                **dict(cfg, **{"hip.skin": skin, "hip.rebuild_blocks": 65528}))
        b = run("boids2d", "f64", params, steps, **dict(cfg, **{"hip.skin": skin, "hip.rebuild": "split"}))
        for t, x, y in zip(steps, a, b):
            assert field_bytes(x) == field_bytes(y), f"skin {skin}: step {t} differs"


def test_forced_spin_limit_is_reported_and_recovers():
    """A phase wait that runs past its spin limit -- impossible at the default
    limit, forced here with a limit of 0 so the first wait gives up --
    abandons the rebuild: the next synchronisation reports ABL_ERR_DEVICE, the
    runtime rebuilds from scratch with the split launches and later runs are
    correct again (equal to the oracle from the state they start from)."""
    from openabl_amd import MEMBERS, AblError, Simulation, load_library, record_dtype
    print("case nlist abandoned one-launch rebuild", flush=True)
    load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    params = boids2d_params(num_agents=4096)
    sim = Simulation("boids2d", "f64", params=params,
                     config={"hip.skin": 1e-9, "hip.barrier_spins": 0})
    sim.init_agents()
    sim.simulate(3)
    with pytest.raises(AblError):
        sim.synchronize()
    dt = record_dtype(MEMBERS["boids2d"], "f64")
    s0 = sim.download(dt)
    sim.simulate(1)
    sim.synchronize()
    out = sim.download(dt)
    sim.close()
    o = Oracle("boids2d").set(**params)
    o.set_agents(s0)
    o.simulate(1)
    want = o.agents()
    for f in ("pos", "velocity"):
        tol = 1e-12 * np.maximum(np.abs(want[f]), 1.0)
        assert (np.abs(out[f] - want[f]) <= tol).all(), f


def test_tile_beyond_16bit_indices_walks_stencil():
    """A tile whose halo holds 65,535 records or more keeps no lists (list
    entries and the lane order hold 16-bit indices) and its agents walk the
    stencil: 66,000 boids in a 40 x 40 box fall in one tile, and the run is
    bit-identical to hip.nlist=0."""
    print("case nlist tile beyond 16-bit indices", flush=True)
    params = boids2d_params(num_agents=66000, agent_density=66000 / 1600.0)
    assert params["max_pos"] == 40.0
    steps = (1, 2)
    a = run("boids2d", "f64", params, steps, **{"hip.cell_scale": 2, "hip.cell_inflate": 1})
    b = run("boids2d", "f64", params, steps, **{"hip.cell_scale": 2, "hip.nlist": 0})
    for t, x, y in zip(steps, a, b):
        assert field_bytes(x) == field_bytes(y), t


def test_two_live_simulations_equal_alone():
    """Two live handles stepping in turn (each on its own stream, so their
    one-launch rebuilds overlap on the device); every timestep rebuilds
    (vanishing skin) and both runs equal the same runs made alone, bit for
    bit."""
    from openabl_amd import MEMBERS, Simulation, load_library, record_dtype
    print("case nlist two live simulations", flush=True)
    cfg = {"hip.cell_scale": 2, "hip.skin": 1e-9}
    pa, pb = boids2d_params(num_agents=20000), boids2d_params(num_agents=12000)
    alone = [run("boids2d", "f64", p, (10,), **cfg)[0] for p in (pa, pb)]
    dt = record_dtype(MEMBERS["boids2d"], "f64")
    lib = load_library()
    lib.abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    a = Simulation("boids2d", "f64", params=pa, config=cfg)
    a.init_agents()
    lib.abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    b = Simulation("boids2d", "f64", params=pb, config=cfg)
    b.init_agents()
    for _ in range(5):
        a.simulate(2)
        b.simulate(2)
    got = [a.download(dt), b.download(dt)]
    a.close()
    b.close()
    for x, y in zip(got, alone):
        assert field_bytes(x) == field_bytes(y)


def test_lists_given_up_when_rebuilt_every_step():
    """With hip.nlist left at its default, an agent type whose lists are
    rebuilt on most timesteps (circle's agents outrun the skin every step)
    goes back to the stencil walk after a sample of the device's rebuild
    count; the runs before, across and after the switch equal the stencil
    walk bit for bit.  With hip.nlist=1 given, the lists stay (same bytes)."""
    print("case nlist given up: circle 65536, 120 steps", flush=True)
    params = circle_params(num_agents=65536, rho=0.1024, r=2.5)
    steps = (20, 60, 120)
    cells = {"hip.cell_inflate": 1.0 + 0.05}  # the list mode's cells in both runs
    a = run("circle", "f64", params, steps, **cells)
    b = run("circle", "f64", params, steps, **dict(cells, **{"hip.nlist": 0}))
    for t, x, y in zip(steps, a, b):
        assert field_bytes(x) == field_bytes(y), t
    c = run("circle", "f64", params, (40,), **dict(cells, **{"hip.nlist": 1}))
    assert field_bytes(c[0]) == field_bytes(run("circle", "f64", params, (40,),
                                                **dict(cells, **{"hip.nlist": 0}))[0])
