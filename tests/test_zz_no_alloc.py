"""abl_hip_reduce_sum_noalloc: the allocation-free sum(Type.member) equals
abl_hip_reduce_sum bit for bit, through the C ABI, on two calls in a row on
one device, the second with a smaller input (the static scratch keeps the
first call's partials), for float, vector, int and bool members."""
import ctypes as C

import numpy as np
import pytest

from orclib import boids2d_params, predator_prey_params

pytestmark = pytest.mark.gpu


def sums(lib, sim, t, member, ncomp):
    a = (C.c_double * ncomp)()
    b = (C.c_double * ncomp)()
    assert lib.abl_hip_reduce_sum(sim.h, t, member.encode(), a) == 0
    assert lib.abl_hip_reduce_sum_noalloc(sim.h, t, member.encode(), b) == 0, \
        lib.abl_hip_last_error().decode()
    return np.array(a[:]), np.array(b[:])


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_sum_equals_allocating_sum(prec):
    from openabl_amd import Simulation, load_library
    lib = load_library()
    print(f"case reduce_sum_noalloc boids2d {prec}", flush=True)
    for n in (1 << 20, 3000):  # the second call reduces fewer blocks
        lib.abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
        sim = Simulation("boids2d", prec, params=boids2d_params(num_agents=n))
        sim.init_agents()
        sim.simulate(2)
        for member in ("pos", "velocity"):
            a, b = sums(lib, sim, 0, member, 2)
            assert a.tobytes() == b.tobytes(), (n, member, a, b)
        sim.close()


def test_noalloc_sum_int_and_bool_members():
    from openabl_amd import Simulation, load_library
    lib = load_library()
    print("case reduce_sum_noalloc predator_prey", flush=True)
    for n in (100000, 2000):
        lib.abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
        p = predator_prey_params(num_agents=n, prey_fraction=0.4, predator_fraction=0.04)
        sim = Simulation("predator_prey", "f64", params=p)
        sim.init_agents()
        sim.simulate(3)
        for t, member in ((2, "avail"), (2, "dead_cycles"), (1, "life"), (0, "pos")):
            nc = 2 if member == "pos" else 1
            a, b = sums(lib, sim, t, member, nc)
            assert a.tobytes() == b.tobytes(), (n, member)
        sim.close()


def test_noalloc_sum_errors():
    from openabl_amd import Simulation, load_library
    lib = load_library()
    sim = Simulation("boids2d", "f64", params=boids2d_params(num_agents=100))
    sim.init_agents()
    out = (C.c_double * 2)()
    assert lib.abl_hip_reduce_sum_noalloc(sim.h, 0, b"nope", out) != 0
    assert lib.abl_hip_reduce_sum_noalloc(sim.h, 5, b"pos", out) != 0
    assert lib.abl_hip_reduce_sum_noalloc(sim.h, 0, b"pos", None) != 0
    sim.close()
