"""FLAME and FlameGPU XML save() (abl_save_flame_xml; asset/c/libabl.c:126-213)
against the reference runtime's own output: tests/golden/save_xml_<prec>.json
holds one record set written by the reference's save() as JSON and as both
XML forms.  The records are rebuilt from the JSON text (6-decimal values
print back to the same text), written through the product's JSON writer
(which must give the JSON text again) and through its XML writer."""
import ctypes as C
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


class TypeInfo(C.Structure):
    _fields_ = [("type", C.c_int), ("offset", C.c_uint), ("name", C.c_char_p), ("is_pos", C.c_bool)]


class AgentInfo(C.Structure):
    _fields_ = [("info", C.POINTER(TypeInfo)), ("offset", C.c_uint), ("name", C.c_char_p)]


class DynArray(C.Structure):
    _fields_ = [("values", C.c_void_p), ("len", C.c_size_t), ("cap", C.c_size_t)]


# type ids of include/abl_hip.h
BOOL, INT, FLOAT, FLOAT2, FLOAT3 = 1, 2, 3, 5, 6


def probe_types(prec):
    f = "<f8" if prec == "f64" else "<f4"
    a = np.dtype([("pos", f, (2,)), ("alive", "u1"), ("life", "<i4"), ("energy", f)], align=True)
    b = np.dtype([("pos", f, (3,)), ("vel", f, (3,))], align=True)
    return (("ProbeA", a, [("pos", FLOAT2, True), ("alive", BOOL, False), ("life", INT, False),
                           ("energy", FLOAT, False)]),
            ("ProbeB", b, [("pos", FLOAT3, True), ("vel", FLOAT3, False)]))


def records(prec, text):
    data = json.loads(text)
    out = []
    for name, dt, members in probe_types(prec):
        rec = np.zeros(len(data[name]), dtype=dt)
        for k, agent in enumerate(data[name]):
            for m, _, _ in members:
                rec[k][m] = agent[m]
        out.append((name, rec, members))
    return out


def save(hip, fn, recs, path, float_bytes, *extra):
    arrs = (DynArray * (2 * len(recs)))()
    ainfo = (AgentInfo * (len(recs) + 1))()
    keep = []
    for t, (name, rec, members) in enumerate(recs):
        infos = (TypeInfo * (len(members) + 1))()
        for k, (m, tid, is_pos) in enumerate(members):
            nb = m.encode()
            keep.append(nb)
            infos[k] = TypeInfo(tid, rec.dtype.fields[m][1], nb, is_pos)
        infos[len(members)] = TypeInfo(0, rec.dtype.itemsize, None, False)
        keep += [infos, rec]
        arrs[2 * t] = DynArray(rec.ctypes.data, len(rec), len(rec))
        nb = name.encode()
        keep.append(nb)
        ainfo[t] = AgentInfo(infos, 2 * t * C.sizeof(DynArray), nb)
    ainfo[len(recs)] = AgentInfo(None, 0, None)
    rc = fn(C.addressof(arrs), C.addressof(ainfo), str(path).encode(), float_bytes, *extra)
    assert rc == 0
    return path.read_text()


@pytest.fixture(scope="module")
def hip():
    from openabl_amd import load_library
    return load_library()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_flame_xml_equals_reference(hip, tmp_path, prec):
    with open(os.path.join(HERE, "golden", f"save_xml_{prec}.json")) as f:
        fx = json.load(f)
    recs = records(prec, fx["json"])
    fb = 8 if prec == "f64" else 4
    assert save(hip, hip.abl_save_json, recs, tmp_path / "s.json", fb) == fx["json"]
    assert save(hip, hip.abl_save_flame_xml, recs, tmp_path / "f.xml", fb, 0) == fx["flame"]
    assert save(hip, hip.abl_save_flame_xml, recs, tmp_path / "g.xml", fb, 1) == fx["flamegpu"]


def test_flamegpu_xml_rejects_scalar_position(hip, tmp_path):
    """FlameGPU positions are vectors: a scalar position member is an error."""
    arrs = (DynArray * 2)()
    infos = (TypeInfo * 2)(TypeInfo(INT, 0, b"life", True), TypeInfo(0, 4, None, False))
    ainfo = (AgentInfo * 2)(AgentInfo(infos, 0, b"ProbeA"), AgentInfo(None, 0, None))
    path = str(tmp_path / "x.xml").encode()
    assert hip.abl_save_flame_xml(C.addressof(arrs), C.addressof(ainfo), path, 8, 1) != 0
    assert hip.abl_save_flame_xml(C.addressof(arrs), C.addressof(ainfo), path, 8, 0) == 0
