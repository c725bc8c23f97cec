"""Launch-shape autotuner for the -b hip runtime.

For one simulation configuration it times a few timesteps under each
candidate runtime configuration -- threads per block ("hip.block"), cells per
granularity ("hip.cell_scale", i.e. cell edge = granularity / scale with a
(2*scale+1)^d stencil), LDS tile shape ("hip.tile", 0 = untiled) and LDS
budget ("hip.lds_slack") -- on a private copy of the state, and returns the
fastest.  Every candidate computes the same neighbour sets (the near test is
exact whatever the cell size), and tile shape and LDS budget do not change
results at all, so tuning never changes results beyond the summation order
of the default cell-ordered visit (which depends on the cell size).

The neighbour-list skin ("hip.skin") is left out of the default space: its
cost is the rebuild rate, which a window of a few timesteps does not see, so
a short search would always pick the smallest skin.  On 1,048,576 boids
over 200 timesteps 0.05 beat 0.02-0.04 by 4-8 %, although the step kernel
alone was faster with the smaller skins; pass "hip.skin" in ``space`` with a
``steps`` window as long as the runs being tuned for.
"""
from __future__ import annotations

import itertools
import time

from .runtime import Simulation


DEFAULT_SPACE = {
    "hip.block": [128, 256],
    "hip.cell_scale": [1, 2, 3],
    "hip.tile": [0, 1, 2, 3],
    "hip.lds_slack": [1.25, 2.0],
}


def _time_config(program, precision, params, records, cfg, steps, warmup):
    sim = Simulation(program, precision, params=params, config=cfg)
    for t, rec in enumerate(records):
        sim.upload(rec, t)
    sim.simulate(warmup)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.simulate(steps)
    sim.synchronize()
    dt = (time.perf_counter() - t0) / steps
    sim.close()
    return dt


def tune(program: str, precision: str, params: dict, records: list, space: dict | None = None,
         steps: int = 10, warmup: int = 3, base_config: dict | None = None):
    """Return (best_config, table) where table lists (config, seconds/step).

    ``records`` holds one numpy record array per agent type (the state every
    candidate starts from)."""
    space = space or DEFAULT_SPACE
    keys = sorted(space)
    table = []
    for values in itertools.product(*(space[k] for k in keys)):
        cfg = dict(base_config or {})
        cfg.update(dict(zip(keys, values)))
        try:
            dt = _time_config(program, precision, params, records, cfg, steps, warmup)
        except Exception as e:  # a configuration the runtime rejects
            table.append((cfg, None, str(e)))
            continue
        table.append((cfg, dt, ""))
    ok = [row for row in table if row[1] is not None]
    best = min(ok, key=lambda row: row[1])[0] if ok else dict(base_config or {})
    return best, table
