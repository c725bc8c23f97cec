"""The launch-shape autotuner (openabl_amd/autotune.py): its search over the
product of the candidate values, its handling of configurations the runtime
rejects (CPU, with the timing replaced), and on the GPU a real search whose
candidates -- tile shapes, LDS budgets, block sizes at one cell scale, walking
stencils -- all compute the same results bit for bit."""
import numpy as np
import pytest

from orclib import boids2d_params, field_bytes


def test_search_covers_the_product_and_skips_rejected(monkeypatch):
    from openabl_amd import autotune
    seen = []

    def fake(program, precision, params, records, cfg, steps, warmup):
        seen.append(dict(cfg))
        if cfg["hip.tile"] == 2:
            raise RuntimeError("rejected")
        # fastest: tile 3 with the larger block
        return 1.0 / (1 + cfg["hip.tile"] + cfg["hip.block"] / 256)

    monkeypatch.setattr(autotune, "_time_config", fake)
    space = {"hip.tile": [0, 1, 2, 3], "hip.block": [128, 256]}
    best, table = autotune.tune("boids2d", "f64", {}, [], space=space,
                                base_config={"hip.nlist": 0})
    assert len(seen) == len(table) == 8
    assert {(c["hip.tile"], c["hip.block"]) for c in seen} == {(t, b) for t in range(4)
                                                            for b in (128, 256)}
    assert all(c["hip.nlist"] == 0 for c in seen)
    assert best == {"hip.nlist": 0, "hip.tile": 3, "hip.block": 256}
    rejected = [row for row in table if row[1] is None]
    assert len(rejected) == 2 and all(row[2] == "rejected" for row in rejected)


def test_search_with_every_candidate_rejected_keeps_the_base(monkeypatch):
    from openabl_amd import autotune

    def fake(*a, **k):
        raise RuntimeError("no")

    monkeypatch.setattr(autotune, "_time_config", fake)
    best, table = autotune.tune("circle", "f64", {}, [], space={"hip.tile": [1, 2]},
                                base_config={"hip.cell_scale": 2})
    assert best == {"hip.cell_scale": 2} and len(table) == 2


@pytest.mark.gpu
def test_tuned_candidates_compute_the_same_results():
    from openabl_amd import MEMBERS, Simulation, load_library, record_dtype
    from openabl_amd.autotune import tune
    print("case autotune boids2d 16384 f64", flush=True)
    params = boids2d_params(num_agents=16384)
    dt = record_dtype(MEMBERS["boids2d"], "f64")
    load_library().abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    sim = Simulation("boids2d", "f64", params=params)
    sim.init_agents()
    recs = [sim.download(dt)]
    sim.close()
    base = {"hip.cell_scale": 2, "hip.nlist": 0}
    space = {"hip.tile": [0, 1, 2, 3], "hip.lds_slack": [1.25, 2.0], "hip.block": [128, 256]}
    best, table = tune("boids2d", "f64", params, recs, space=space, steps=2, warmup=1,
                       base_config=base)
    assert all(row[1] is not None for row in table), [row for row in table if row[1] is None]
    assert best in [row[0] for row in table]
    outs = []
    for cfg, _, _ in table:
        s = Simulation("boids2d", "f64", params=params, config=cfg)
        s.upload(recs[0], 0)
        s.simulate(3)
        outs.append(field_bytes(s.download(dt)))
        s.close()
    assert all(o == outs[0] for o in outs[1:])
