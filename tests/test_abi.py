"""CPU tests of the C ABI library (no GPU call is made).

* libabl_hip.so loads and exports every function include/abl_hip.h declares;
* its host runtime -- the generator and save() that replace asset/c/libabl.c
  for -b hip -- reproduces the reference's draws and JSON bytes (pinned by
  the golden fixtures written from the reference's own libabl.c);
* the Python mirror lays agent records out exactly as the generated C
  structs (checked against the oracle's RTTI).
"""
import ctypes as C
import json
import os
import re
import subprocess

import numpy as np
import pytest

from orclib import Oracle, circle_params, game_of_life_params, boids2d_params

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "abl_hip.h")
GOLDEN = os.path.join(REPO, "tests", "golden")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(abl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function(hip):
    names = header_functions()
    assert len(names) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", hip._name], capture_output=True,
                         text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if " T " in line)
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_python_binding_covers_header(hip):
    from openabl_amd import runtime
    assert sorted(runtime._SIGNATURES) == header_functions()


def test_programs_listed(hip):
    names = [hip.abl_hip_program_name(i).decode() for i in range(hip.abl_hip_program_count())]
    assert names == ["circle", "circle3d", "boids2d", "game_of_life", "predator_prey"]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_host_generator_matches_reference(hip, prec):
    """abl_random_float(f)/abl_random_int reproduce asset/c/libabl.c:8-39."""
    cases = json.load(open(os.path.join(GOLDEN, f"rng_{prec}.json")))
    rf = hip.abl_random_float if prec == "f64" else hip.abl_random_floatf
    for case in cases:
        hip.abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)
        if case["kind"] == "float":
            got = [float(rf(case["lo"], case["hi"])).hex() for _ in range(case["count"])]
            want = [float.fromhex(v).hex() for v in case["values"]]
        elif case["kind"] == "int":
            got = [hip.abl_random_int(case["lo"], case["hi"]) for _ in range(case["count"])]
            want = [int(v) for v in case["values"]]
        else:
            got, want = [], []
            for line in case["values"]:
                f, a, b = line.split()
                # the probe's printf arguments are evaluated right to left by gcc
                i2 = hip.abl_random_int(5, 12)
                i1 = hip.abl_random_int(0, 9)
                got.append((float(rf(-3, 7)).hex(), i1, i2))
                want.append((float.fromhex(f).hex(), int(a), int(b)))
        assert got == want, case["kind"]
    hip.abl_rng_seed(0xDEADBEEF, 0xBEEFDEAD)


class TypeInfo(C.Structure):
    _fields_ = [("type", C.c_int), ("offset", C.c_uint), ("name", C.c_char_p), ("is_pos", C.c_bool)]


class AgentInfo(C.Structure):
    _fields_ = [("info", C.POINTER(TypeInfo)), ("offset", C.c_uint), ("name", C.c_char_p)]


class DynArray(C.Structure):
    _fields_ = [("values", C.c_void_p), ("len", C.c_size_t), ("cap", C.c_size_t)]


KIND = {"u1": 1, "<i4": 2}


def save_with_product(hip, records, type_name, path, float_bytes):
    """Call abl_save_json on one agent type exactly as generated code would:
    a struct of dyn_arrays plus RTTI tables."""
    dt = records.dtype
    infos = (TypeInfo * (len(dt.names) + 1))()
    keep = []
    for k, name in enumerate(dt.names):
        sub, off = dt.fields[name][0], dt.fields[name][1]
        if sub.shape == (2,):
            t = 5
        elif sub.shape == (3,):
            t = 6
        elif sub.kind == "f":
            t = 3
        else:
            t = KIND[sub.str if sub.str != "|u1" else "u1"]
        nb = name.encode()
        keep.append(nb)
        infos[k] = TypeInfo(t, off, nb, name == "pos")
    infos[len(dt.names)] = TypeInfo(0, dt.itemsize, None, False)
    arrs = (DynArray * 2)()
    buf = np.ascontiguousarray(records)
    arrs[0] = DynArray(buf.ctypes.data, len(buf), len(buf))
    ainfo = (AgentInfo * 2)()
    tn = type_name.encode()
    ainfo[0] = AgentInfo(infos, 0, tn)
    ainfo[1] = AgentInfo(None, 0, None)
    rc = hip.abl_save_json(C.addressof(arrs), C.addressof(ainfo), str(path).encode(), float_bytes)
    assert rc == 0


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("model,params,tname", [
    ("circle", circle_params(num_agents=300), "Point"),
    ("boids2d", boids2d_params(num_agents=300), "Boid"),
    ("game_of_life", game_of_life_params(num_agents=400, alive_fraction=0.4), "Cell"),
])
def test_host_save_matches_oracle_save(hip, tmp_path, prec, model, params, tname):
    """abl_save_json writes the bytes of save() (asset/c/libabl.c:46-124) --
    the oracle's save is itself pinned by the reference-runtime fixtures."""
    o = Oracle(model, prec).set(**params).init()
    o.simulate(2)
    recs = o.agents()
    o.save(str(tmp_path / "orc.json"))
    save_with_product(hip, recs, tname, tmp_path / "hip.json", 8 if prec == "f64" else 4)
    assert (tmp_path / "hip.json").read_bytes() == (tmp_path / "orc.json").read_bytes()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_record_layout_matches_generated_structs(prec):
    """The Python mirror's record dtypes equal the generated C structs."""
    from openabl_amd import MEMBERS, record_dtype
    for model in ("circle", "circle3d", "boids2d", "game_of_life"):
        o = Oracle(model, prec)
        assert record_dtype(MEMBERS[model], prec) == o.dtypes[0], model


def test_save_reports_unwritable_path(hip, tmp_path):
    rec = np.zeros(2, dtype=[("pos", "<f8", 2)])
    arrs = (DynArray * 1)()
    arrs[0] = DynArray(rec.ctypes.data, 2, 2)
    infos = (TypeInfo * 2)(TypeInfo(5, 0, b"pos", True), TypeInfo(0, 16, None, False))
    ainfo = (AgentInfo * 2)(AgentInfo(infos, 0, b"Point"), AgentInfo(None, 0, None))
    rc = hip.abl_save_json(C.addressof(arrs), C.addressof(ainfo),
                           str(tmp_path / "no" / "dir" / "x.json").encode(), 8)
    assert rc == 1


def test_derived_constants_fold_like_reference():
    """-P values re-fold derived globals at 6 significant digits
    (src/AnalysisVisitor.cpp:403-472, GenericPrinter.cpp:135-146)."""
    from openabl_amd import derived_params
    assert derived_params("circle", {"num_agents": 1024, "rho": 0.1024})["W"] == 100.0
    assert derived_params("boids2d", {"num_agents": 1048576, "agent_density": 0.065536})["max_pos"] == 4000.0
    assert derived_params("circle3d", {"num_agents": 1048576, "rho": 1048576 / 1e9})["W"] == 1000.0
    assert derived_params("circle", {"num_agents": 1000, "rho": 0.05})["W"] == 141.421
    assert derived_params("game_of_life", {"num_agents": 1048576})["size"] == 1024
