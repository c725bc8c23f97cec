"""Python binding of the -b hip runtime's C ABI (include/abl_hip.h).

This mirrors, for Python callers, what a program generated for ``-b hip``
does in C++: create a simulation of a compiled step program, apply ``-P``
parameters (derived constants folded like the reference compiler,
``src/AnalysisVisitor.cpp:403-472``), run the initialisation of ``main()``,
``simulate(T)``, read agents back in the reference's array order and write
the same JSON ``save()`` does (``asset/c/libabl.c:46-124``).

There is no fallback: if ``openabl_amd/lib/libabl_hip.so`` is missing the
import fails, and every runtime error raises :class:`AblError` with the
library's message.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libabl_hip.so")

PREC = {"f64": 0, "f32": 1, "f32c": 2}
ABL_OK = 0
ERR_NAMES = {1: "ABL_ERR_ARG", 2: "ABL_ERR_STATE", 3: "ABL_ERR_DEVICE", 4: "ABL_ERR_MEMORY",
             5: "ABL_ERR_UNSUPPORTED"}

# Every entry point include/abl_hip.h declares, with its ctypes signature.
_SIGNATURES = {
    "abl_hip_last_error": ([], C.c_char_p),
    "abl_random_float": ([C.c_double, C.c_double], C.c_double),
    "abl_random_floatf": ([C.c_float, C.c_float], C.c_float),
    "abl_random_int": ([C.c_int, C.c_int], C.c_int),
    "abl_rng_seed": ([C.c_uint64, C.c_uint64], None),
    "abl_rng_state": ([C.c_void_p], None),
    "abl_save_json": ([C.c_void_p, C.c_void_p, C.c_char_p, C.c_int], C.c_int),
    "abl_save_flame_xml": ([C.c_void_p, C.c_void_p, C.c_char_p, C.c_int, C.c_int], C.c_int),
    "abl_hip_program_count": ([], C.c_int),
    "abl_hip_program_name": ([C.c_int], C.c_char_p),
    "abl_hip_create": ([C.c_char_p, C.c_int, C.c_void_p], C.c_int),
    "abl_hip_destroy": ([C.c_void_p], C.c_int),
    "abl_hip_set_param": ([C.c_void_p, C.c_char_p, C.c_double], C.c_int),
    "abl_hip_get_param": ([C.c_void_p, C.c_char_p, C.c_void_p], C.c_int),
    "abl_hip_set_config": ([C.c_void_p, C.c_char_p, C.c_char_p], C.c_int),
    "abl_hip_agent_type_count": ([C.c_void_p], C.c_int),
    "abl_hip_agent_type_name": ([C.c_void_p, C.c_int], C.c_char_p),
    "abl_hip_agent_size": ([C.c_void_p, C.c_int], C.c_size_t),
    "abl_hip_upload": ([C.c_void_p, C.c_int, C.c_void_p, C.c_size_t], C.c_int),
    "abl_hip_download": ([C.c_void_p, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p], C.c_int),
    "abl_hip_count": ([C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "abl_hip_init_agents": ([C.c_void_p], C.c_int),
    "abl_hip_upload_all": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "abl_hip_download_all": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "abl_hip_simulate": ([C.c_void_p, C.c_int], C.c_int),
    "abl_hip_synchronize": ([C.c_void_p], C.c_int),
    "abl_hip_neighbours": ([C.c_void_p, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p], C.c_int),
    "abl_hip_footprint": ([C.c_void_p, C.c_void_p], C.c_int),
    "abl_hip_last_timing": ([C.c_void_p, C.c_void_p], C.c_int),
    "abl_hip_download_ids": ([C.c_void_p, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p], C.c_int),
    "abl_hip_reduce_sum": ([C.c_void_p, C.c_int, C.c_char_p, C.c_void_p], C.c_int),
    "abl_hip_reduce_sum_noalloc": ([C.c_void_p, C.c_int, C.c_char_p, C.c_void_p], C.c_int),
    "abl_hip_reduce_count_eq": ([C.c_void_p, C.c_int, C.c_char_p, C.c_double, C.c_void_p], C.c_int),
    "abl_hip_last_exec_time": ([C.c_void_p, C.c_void_p], C.c_int),
    "abl_hip_log_csv": ([C.c_void_p, C.c_char_p], C.c_int),
    "abl_hip_log_text": ([C.c_void_p, C.c_char_p, C.c_size_t, C.c_void_p], C.c_int),
    "abl_hip_query": ([C.c_void_p, C.c_void_p], C.c_int),
    "abl_hip_estimate_footprint": ([C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                    C.c_size_t, C.c_size_t, C.c_void_p], C.c_int),
    "abl_server_last_error": ([], C.c_char_p),
    "abl_server_create": ([C.c_size_t, C.c_int, C.c_int, C.c_void_p], C.c_int),
    "abl_server_submit": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "abl_server_status": ([C.c_void_p, C.c_uint64, C.c_void_p], C.c_int),
    "abl_server_wait": ([C.c_void_p, C.c_uint64], C.c_int),
    "abl_server_result": ([C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p], C.c_int),
    "abl_server_run_log": ([C.c_void_p, C.c_uint64, C.c_char_p, C.c_size_t, C.c_void_p], C.c_int),
    "abl_server_run_times": ([C.c_void_p, C.c_uint64, C.c_void_p], C.c_int),
    "abl_server_stats": ([C.c_void_p, C.c_void_p], C.c_int),
    "abl_server_destroy": ([C.c_void_p], C.c_int),
    "abl_ens_last_error": ([], C.c_char_p),
    "abl_ens_member_seed": ([C.c_uint64, C.c_uint64, C.c_int, C.c_void_p], C.c_int),
    "abl_ens_create": ([C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                        C.c_void_p], C.c_int),
    "abl_ens_set_config": ([C.c_void_p, C.c_char_p, C.c_char_p], C.c_int),
    "abl_ens_init": ([C.c_void_p, C.c_uint64, C.c_uint64], C.c_int),
    "abl_ens_simulate": ([C.c_void_p, C.c_int], C.c_int),
    "abl_ens_synchronize": ([C.c_void_p, C.c_void_p], C.c_int),
    "abl_ens_member_count": ([C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "abl_ens_agents": ([C.c_void_p, C.c_void_p], C.c_int),
    "abl_ens_download": ([C.c_void_p, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p], C.c_int),
    "abl_ens_destroy": ([C.c_void_p], C.c_int),
}


class AblError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libabl_hip.so (no GPU call is made by loading)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run __graft_entry__.build() "
                          "or make -C openabl_amd/csrc")
    lib = C.CDLL(path)
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _check(rc: int):
    if rc != ABL_OK:
        raise AblError(rc, load_library().abl_hip_last_error().decode())


def fold6(x: float) -> float:
    """A folded float constant as the reference prints it (6 significant
    digits, ``src/backend/GenericPrinter.cpp:135-146``)."""
    return float("%.6g" % x)


# Parameters of each program with the derived constants the reference
# compiler folds from them (src/AnalysisVisitor.cpp:403-472).
def derived_params(program: str, params: dict) -> dict:
    p = dict(params)
    if program == "circle":
        if "num_agents" in p or "rho" in p:
            p["W"] = fold6(math.sqrt(p.get("num_agents", 1000) / p.get("rho", 0.05)))
    elif program == "circle3d":
        if "num_agents" in p or "rho" in p:
            p["W"] = fold6(float(np.cbrt(p.get("num_agents", 1000) / p.get("rho", 0.05))))
    elif program == "boids2d":
        if "num_agents" in p or "agent_density" in p:
            p["max_pos"] = fold6(math.sqrt(p.get("num_agents", 65536) / p.get("agent_density", 0.065536)))
    elif program == "game_of_life":
        if "num_agents" in p:
            p["size"] = int(math.sqrt(p["num_agents"]))
    elif program == "predator_prey":
        sm = p.get("SPACE_MULT", 1.0)
        if "SPACE_MULT" in p:
            for name, base in (("PRED_PREY_INTERACTION_RADIUS", 0.1),
                               ("PREY_GROUP_COHESION_RADIUS", 0.12),
                               ("SAME_SPECIES_AVOIDANCE_RADIUS", 0.035),
                               ("GRASS_EAT_DISTANCE", 0.02), ("PRED_KILL_DISTANCE", 0.02)):
                p[name] = fold6(base * sm)
        if "num_agents" in p or "agent_density" in p or "SPACE_MULT" in p:
            p["env_size"] = fold6(math.sqrt(p.get("num_agents", 3200) / p.get("agent_density", 1600.0)) * sm)
    return p


_FIELD = {1: "u1", 2: "<i4"}


class Simulation:
    """One simulation of a compiled step program on the current device."""

    def __init__(self, program: str, precision: str = "f64", params: dict | None = None,
                 config: dict | None = None):
        lib = load_library()
        self.lib = lib
        self.program = program
        self.precision = precision
        h = C.c_void_p()
        _check(lib.abl_hip_create(program.encode(), PREC[precision], C.byref(h)))
        self.h = h
        for k, v in derived_params(program, params or {}).items():
            self.set_param(k, v)
        for k, v in (config or {}).items():
            self.set_config(k, v)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.abl_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- parameters / config ---------------------------------------------------
    def set_param(self, name: str, value: float):
        _check(self.lib.abl_hip_set_param(self.h, name.encode(), float(value)))

    def get_param(self, name: str) -> float:
        v = C.c_double()
        _check(self.lib.abl_hip_get_param(self.h, name.encode(), C.byref(v)))
        return v.value

    def set_config(self, key: str, value):
        _check(self.lib.abl_hip_set_config(self.h, key.encode(), str(value).encode()))

    # -- agents -------------------------------------------------------------------
    @property
    def num_types(self) -> int:
        return self.lib.abl_hip_agent_type_count(self.h)

    def record_size(self, t: int = 0) -> int:
        return self.lib.abl_hip_agent_size(self.h, t)

    def init_agents(self):
        _check(self.lib.abl_hip_init_agents(self.h))

    def upload(self, records: np.ndarray, t: int = 0):
        arr = np.ascontiguousarray(records)
        if arr.dtype.itemsize != self.record_size(t):
            raise ValueError(f"record size {arr.dtype.itemsize} != {self.record_size(t)}")
        _check(self.lib.abl_hip_upload(self.h, t, arr.ctypes.data, len(arr)))

    def count(self, t: int = 0) -> int:
        n = C.c_size_t()
        _check(self.lib.abl_hip_count(self.h, t, C.byref(n)))
        return n.value

    def download(self, dtype: np.dtype, t: int = 0) -> np.ndarray:
        n = self.count(t)
        out = np.zeros(n, dtype=dtype)
        if dtype.itemsize != self.record_size(t):
            raise ValueError(f"record size {dtype.itemsize} != {self.record_size(t)}")
        got = C.c_size_t()
        _check(self.lib.abl_hip_download(self.h, t, out.ctypes.data if n else None, n, C.byref(got)))
        return out[: got.value]

    # -- running ------------------------------------------------------------------
    def simulate(self, timesteps: int):
        _check(self.lib.abl_hip_simulate(self.h, int(timesteps)))

    def synchronize(self):
        _check(self.lib.abl_hip_synchronize(self.h))

    def neighbours(self, step: int = 0):
        n = C.c_size_t()
        _check(self.lib.abl_hip_neighbours(self.h, step, None, 0, C.byref(n)))
        out = np.zeros(2 * max(n.value, 1), dtype=np.int64)
        got = C.c_size_t()
        _check(self.lib.abl_hip_neighbours(self.h, step, out.ctypes.data, n.value, C.byref(got)))
        pairs = out[: 2 * got.value].reshape(-1, 2)
        return pairs[:, 0].copy(), pairs[:, 1].copy()

    def ids(self, t: int = 0) -> np.ndarray:
        """Ids of type t's agents in download order."""
        n = self.count(t)
        out = np.zeros(max(n, 1), dtype=np.int64)
        got = C.c_size_t()
        _check(self.lib.abl_hip_download_ids(self.h, t, out.ctypes.data, n, C.byref(got)))
        return out[: got.value]

    def reduce_sum(self, t: int, member: str, ncomp: int = 1):
        out = (C.c_double * ncomp)()
        _check(self.lib.abl_hip_reduce_sum(self.h, t, member.encode(), out))
        return list(out) if ncomp > 1 else out[0]

    def count_eq(self, t: int, member: str, value) -> int:
        n = C.c_size_t()
        _check(self.lib.abl_hip_reduce_count_eq(self.h, t, member.encode(), float(value), C.byref(n)))
        return n.value

    def log_text(self) -> str:
        n = C.c_size_t()
        _check(self.lib.abl_hip_log_text(self.h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        _check(self.lib.abl_hip_log_text(self.h, buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    def footprint(self) -> int:
        b = C.c_size_t()
        _check(self.lib.abl_hip_footprint(self.h, C.byref(b)))
        return b.value

    def last_timing(self):
        out = (C.c_double * 3)()
        _check(self.lib.abl_hip_last_timing(self.h, out))
        return list(out)


def record_dtype(members, precision: str = "f64") -> np.dtype:
    """numpy dtype of a generated agent struct: members [(name, kind)] with
    kind 'bool', 'int', 'float', 'float2', 'float3', laid out by the C rules
    the C backend's struct follows (src/backend/CPrinter.cpp:250-253)."""
    fb = 4 if precision == "f32" else 8  # f32c keeps the FP64 struct layout
    fcode = "<f8" if fb == 8 else "<f4"
    names, formats, offsets = [], [], []
    off, amax = 0, 1
    for name, kind in members:
        size, fmt = {"bool": (1, "u1"), "int": (4, "<i4"), "float": (fb, fcode),
                     "float2": (fb, (fcode, 2)), "float3": (fb, (fcode, 3))}[kind]
        ncomp = {"float2": 2, "float3": 3}.get(kind, 1)
        off = (off + size - 1) // size * size
        names.append(name)
        formats.append(fmt)
        offsets.append(off)
        off += size * ncomp
        amax = max(amax, size)
    itemsize = (off + amax - 1) // amax * amax
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": itemsize})


_MOVER = [("pos", "float2"), ("dir", "float2"), ("steer", "float2"), ("life", "int")]

# Agent types of each program, in declaration order, with their members.
AGENTS = {
    "circle": [("Point", [("pos", "float2")])],
    "circle3d": [("Point", [("pos", "float3")])],
    "boids2d": [("Boid", [("pos", "float2"), ("velocity", "float2")])],
    "game_of_life": [("Cell", [("pos", "float2"), ("alive", "bool")])],
    "predator_prey": [("Predator", _MOVER), ("Prey", _MOVER),
                      ("Grass", [("pos", "float2"), ("dead_cycles", "int"), ("avail", "bool")])],
}
MEMBERS = {k: v[0][1] for k, v in AGENTS.items()}
AGENT_NAMES = {k: v[0][0] for k, v in AGENTS.items()}
