"""End to end against the reference runtime: the HIP path's saved JSON.

tests/golden/runs_<prec>.json hold the SHA-256 of the file save() writes at
the end of small runs of each model, built from the restated generated
programs linked against the reference's own asset/c/libabl.c
(tests/golden/make_golden.py).  Here the same runs go through the C ABI the
way a generated -b hip program drives it: host initialisation, simulate in
the reference's index order (hip.order=index), abl_hip_download_all into the
generated agent_struct's dyn_arrays, then abl_save_json -- and the file's
SHA-256 must equal the fixture's.  This ties the device path to the
reference runtime's bytes, not only to the oracle.
"""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TYPE_ID = {"bool": 1, "int": 2, "float": 3, "float2": 5, "float3": 6}


class DynArray(C.Structure):
    _fields_ = [("values", C.c_void_p), ("len", C.c_size_t), ("cap", C.c_size_t)]


class TypeInfo(C.Structure):
    _fields_ = [("type", C.c_int), ("offset", C.c_uint), ("name", C.c_char_p),
                ("is_pos", C.c_bool)]


class AgentInfo(C.Structure):
    _fields_ = [("info", C.POINTER(TypeInfo)), ("offset", C.c_uint), ("name", C.c_char_p)]


def generated_tables(program, precision):
    """The agent_struct, its RTTI and agents_info table a generated program
    holds (src/backend/CPrinter.cpp:250-302): per type a dyn_array and its
    _dbuf, members in declaration order."""
    from openabl_amd import AGENTS, record_dtype
    types = AGENTS[program]
    Struct = type("agent_struct", (C.Structure,),
                  {"_fields_": [(f"a{k}_{s}", DynArray) for k in range(len(types))
                                for s in ("buf", "dbuf")]})
    keep = []
    infos = (AgentInfo * (len(types) + 1))()
    for k, (name, members) in enumerate(types):
        dt = record_dtype(members, precision)
        ti = (TypeInfo * (len(members) + 1))()
        for m, (mname, kind) in enumerate(members):
            bname = mname.encode()
            keep.append(bname)
            ti[m] = TypeInfo(TYPE_ID[kind], dt.fields[mname][1], bname, mname == "pos")
        ti[len(members)] = TypeInfo(0, dt.itemsize, None, False)
        keep.append(ti)
        bn = name.encode()
        keep.append(bn)
        infos[k] = AgentInfo(ti, getattr(Struct, f"a{k}_buf").offset, bn)
    infos[len(types)] = AgentInfo(None, 0, None)
    return Struct(), infos, keep


def cases(prec):
    with open(os.path.join(HERE, "golden", f"runs_{prec}.json")) as f:
        return [(prec, r) for r in json.load(f)]


@pytest.mark.parametrize("prec,case", cases("f64") + cases("f32"),
                         ids=lambda v: v if isinstance(v, str) else
                         f"{v['model']}-{v['params']['num_agents']}")
def test_saved_json_equals_reference_runtime(prec, case, tmp_path):
    from openabl_amd import Simulation, load_library
    lib = load_library()
    model, params = case["model"], case["params"]
    steps = int(params["num_timesteps"])
    print(f"case golden save {model} {prec} n={params['num_agents']} steps={steps}", flush=True)
    lib.abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    sim = Simulation(model, prec, params=params, config={"hip.order": "index"})
    sim.init_agents()
    sim.simulate(steps)
    agents, info, keep = generated_tables(model, prec)
    rc = lib.abl_hip_download_all(sim.h, C.byref(agents), info)
    assert rc == 0, lib.abl_hip_last_error().decode()
    sim.close()
    out = tmp_path / "out.json"
    assert lib.abl_save_json(C.byref(agents), info, str(out).encode(),
                             4 if prec == "f32" else 8) == 0
    data = out.read_bytes()
    assert len(data) == case["bytes"]
    assert hashlib.sha256(data).hexdigest() == case["sha256"]


def test_upload_all_round_trip():
    """upload_all takes the generated agent_struct back: download_all,
    upload_all into a fresh simulation, download_all again -- the same
    records, and the struct-size check rejects a mismatched table."""
    from openabl_amd import AGENTS, Simulation, load_library, record_dtype
    lib = load_library()
    lib.abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
    p = {"num_agents": 20000, "prey_fraction": 0.4, "predator_fraction": 0.04}
    sim = Simulation("predator_prey", "f64", params=p)
    sim.init_agents()
    sim.simulate(3)
    a, info, keep = generated_tables("predator_prey", "f64")
    assert lib.abl_hip_download_all(sim.h, C.byref(a), info) == 0
    b2 = Simulation("predator_prey", "f64", params=p)
    assert lib.abl_hip_upload_all(b2.h, C.byref(a), info) == 0
    c, info2, keep2 = generated_tables("predator_prey", "f64")
    assert lib.abl_hip_download_all(b2.h, C.byref(c), info2) == 0
    for k, (name, members) in enumerate(AGENTS["predator_prey"]):
        dt = record_dtype(members, "f64")
        da, dc = getattr(a, f"a{k}_buf"), getattr(c, f"a{k}_buf")
        assert da.len == dc.len and da.len > 0
        x = np.ctypeslib.as_array((C.c_char * (da.len * dt.itemsize)).from_address(da.values))
        y = np.ctypeslib.as_array((C.c_char * (dc.len * dt.itemsize)).from_address(dc.values))
        assert bytes(x) == bytes(y), name
    # a table whose struct size disagrees with the program's is rejected
    bad, binfo, bkeep = generated_tables("predator_prey", "f32")
    assert lib.abl_hip_upload_all(b2.h, C.byref(bad), binfo) != 0
    sim.close()
    b2.close()
