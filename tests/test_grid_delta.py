"""The delta re-sort (hip.grid=delta, the default: after a step only the
agents that changed cell touch atomics) gives the same cell-sorted order as
the full counting sort (hip.grid=full), so whole runs are bit-identical, in
steady state and when most agents change cell every step."""
import numpy as np
import pytest

from orclib import boids2d_params, circle3d_params, circle_params, field_bytes

pytestmark = pytest.mark.gpu


def run(model, prec, params, steps, grid, **cfg):
    from openabl_amd import MEMBERS, Simulation, load_library, record_dtype
    load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    sim = Simulation(model, prec, params=params, config=dict(cfg, **{"hip.grid": grid}))
    sim.init_agents()
    out = []
    done = 0
    for upto in steps:
        sim.simulate(upto - done)
        done = upto
        out.append((sim.download(record_dtype(MEMBERS[model], prec)), sim.neighbours(0)))
    sim.close()
    return out


CASES = [
    ("boids2d", "f64", boids2d_params(num_agents=20000), {}),
    ("boids2d", "f32", boids2d_params(num_agents=20000), {"hip.cell_scale": 2}),
    ("boids2d", "f32c", boids2d_params(num_agents=20000), {"hip.cell_scale": 3}),
    ("circle", "f64", circle_params(num_agents=4096), {}),
    ("circle", "f64", circle_params(num_agents=4096), {"hip.order": "index"}),
    # dense clusters: agents cross cells every step in the first steps
    ("circle3d", "f64", circle3d_params(num_agents=32768, rho=32768 / 1e9, r=4.0,
                                        num_clusters=2, cluster_sigma=20.0), {}),
    ("circle3d", "f32c", circle3d_params(num_agents=32768, rho=32768 / 1e9, r=4.0,
                                         num_clusters=2, cluster_sigma=20.0), {}),
]


@pytest.mark.parametrize("model,prec,params,cfg", CASES,
                         ids=[f"{c[0]}-{c[1]}-{i}" for i, c in enumerate(CASES)])
def test_delta_equals_full(model, prec, params, cfg):
    print(f"case grid delta {model} {prec} {cfg}", flush=True)
    steps = (1, 2, 5, 20)
    a = run(model, prec, params, steps, "delta", **cfg)
    b = run(model, prec, params, steps, "full", **cfg)
    for t, (x, y) in zip(steps, zip(a, b)):
        assert field_bytes(x[0]) == field_bytes(y[0]), t
        assert np.array_equal(x[1][0], y[1][0]) and np.array_equal(x[1][1], y[1][1]), t


def test_delta_after_upload_and_param_change():
    """A new upload or a changed grid (parameter, cell scale) falls back to the
    full sort; results stay identical to full-sort runs."""
    from openabl_amd import MEMBERS, Simulation, record_dtype
    params = boids2d_params(num_agents=5000)
    dt = record_dtype(MEMBERS["boids2d"], "f64")
    rng = np.random.default_rng(4)
    recs = np.zeros(5000, dtype=dt)
    recs["pos"] = rng.uniform(0, params["max_pos"], (5000, 2))
    recs["velocity"] = rng.uniform(-1, 1, (5000, 2))
    outs = []
    for grid in ("delta", "full"):
        sim = Simulation("boids2d", "f64", params=params, config={"hip.grid": grid})
        sim.upload(recs)
        sim.simulate(3)
        sim.upload(sim.download(dt)[::-1].copy())
        sim.simulate(3)
        sim.set_config("hip.cell_scale", 2)
        sim.simulate(3)
        sim.set_param("interaction_radius", 7.5)
        sim.simulate(3)
        outs.append(sim.download(dt))
        sim.close()
    assert field_bytes(outs[0]) == field_bytes(outs[1])


TILE_CASES = [
    ("boids2d", boids2d_params(num_agents=20000), 1),
    ("boids2d", boids2d_params(num_agents=20000), 2),
    ("circle", circle_params(num_agents=4096), 1),
    ("circle3d", circle3d_params(num_agents=32768, rho=32768 / 1e9, r=4.0, num_clusters=2,
                                 cluster_sigma=20.0), 1),
]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("model,params,cs", TILE_CASES,
                         ids=[f"{c[0]}-cs{c[2]}-{i}" for i, c in enumerate(TILE_CASES)])
def test_tiled_kernels_equal_untiled(model, params, cs, prec):
    """The LDS-tiled neighbour kernels (tile shapes 1-3, any LDS budget) visit
    the accepted neighbours in the same order as the untiled kernel (row by
    row of the stencil, cell-sorted order inside a row; narrowing only skips
    cells that cannot hold a neighbour), so runs are bit-identical."""
    print(f"case tiles {model} {prec} cs={cs}", flush=True)
    ref = run(model, prec, params, (1, 5), "delta", **{"hip.tile": 0, "hip.cell_scale": cs})
    for tile in (1, 2, 3):
        for slack in (2.0, 1.0):
            got = run(model, prec, params, (1, 5), "delta",
                      **{"hip.tile": tile, "hip.cell_scale": cs, "hip.lds_slack": slack,
                         "hip.nlist": 0})
            for (x, _), (y, _) in zip(got, ref):
                assert field_bytes(x) == field_bytes(y), (tile, slack)
