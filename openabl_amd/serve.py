"""Python binding of the serving engine (abl_server_*, include/abl_hip.h).

A :class:`Server` accepts simulation runs of any compiled program while
others are on the GPU, admits them by device-memory footprint, overlaps their
work on the device and reports per-run latency and p50/p99 over the served
mix.  The engine itself is native (openabl_amd/csrc/abl_serve.cpp).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from .runtime import AGENTS, AblError, derived_params, load_library, record_dtype

REF_RNG = (0xDEADBEEF, 0xBEEFDEAD)


class RunRequest(C.Structure):
    _fields_ = [("program", C.c_char_p), ("precision", C.c_int), ("num_params", C.c_int),
                ("param_names", C.POINTER(C.c_char_p)), ("param_values", C.POINTER(C.c_double)),
                ("timesteps", C.c_int), ("rng_state", C.c_uint64 * 2),
                ("agents_hint", C.c_size_t), ("cells_hint", C.c_size_t),
                ("num_config", C.c_int), ("config_keys", C.POINTER(C.c_char_p)),
                ("config_values", C.POINTER(C.c_char_p))]


def _check(rc: int):
    if rc != 0:
        raise AblError(rc, load_library().abl_server_last_error().decode())


def sizing(program: str, params: dict) -> tuple[int, int]:
    """(agents, grid cells) a run of `program` with these parameters holds,
    for admission; predator_prey is given room to double."""
    p = derived_params(program, params)
    n = int(p.get("num_agents", 1000))
    if program == "circle":
        side, g = p.get("W", 141.421), 2 * p.get("r", 5.0)
        cells = math.ceil(side / g) ** 2
    elif program == "circle3d":
        side, g = p.get("W", 27.1442), 2 * p.get("r", 5.0)
        cells = math.ceil(side / g) ** 3
    elif program == "boids2d":
        side, g = p.get("max_pos", 1000.0), p.get("interaction_radius", 10.0)
        cells = math.ceil(side / g) ** 2
    elif program == "game_of_life":
        size = int(p.get("size", 20))
        n = size * size
        cells = math.ceil(size / 1.5) ** 2
    else:  # predator_prey
        side = p.get("env_size", 1.41421)
        cells = math.ceil(side / p.get("PREY_GROUP_COHESION_RADIUS", 0.12)) ** 2
        n = 2 * n
    return n, int(cells) * 4  # up to hip.cell_scale 2 in 2-D


def estimate_footprint(program: str, precision: str, params: dict, agents: int,
                       cells: int = 0) -> int:
    """Upper bound of the device bytes a run of `program` with these
    parameters holds while it has at most `agents` agents per type
    (abl_hip_estimate_footprint)."""
    p = derived_params(program, params or {})
    names = [k.encode() for k in p]
    arr_n = (C.c_char_p * max(len(names), 1))(*names)
    arr_v = (C.c_double * max(len(p), 1))(*[float(v) for v in p.values()])
    out = C.c_size_t()
    lib = load_library()
    rc = lib.abl_hip_estimate_footprint(program.encode(), {"f64": 0, "f32": 1, "f32c": 2}[precision],
                                        len(names), arr_n, arr_v, int(agents), int(cells),
                                        C.byref(out))
    if rc != 0:
        raise AblError(rc, lib.abl_hip_last_error().decode())
    return out.value


class Server:
    def __init__(self, hbm_budget: int, max_active: int = 8, chunk_steps: int = 20):
        self.lib = load_library()
        h = C.c_void_p()
        _check(self.lib.abl_server_create(int(hbm_budget), max_active, chunk_steps, C.byref(h)))
        self.h = h
        self._keep = {}

    def submit(self, program: str, precision: str = "f64", params: dict | None = None,
               timesteps: int = 100, rng_state=REF_RNG, config: dict | None = None) -> int:
        p = derived_params(program, params or {})
        names = [k.encode() for k in p]
        arr_n = (C.c_char_p * max(len(names), 1))(*names)
        arr_v = (C.c_double * max(len(p), 1))(*[float(v) for v in p.values()])
        cfg = config or {}
        arr_k = (C.c_char_p * max(len(cfg), 1))(*[k.encode() for k in cfg])
        arr_c = (C.c_char_p * max(len(cfg), 1))(*[str(v).encode() for v in cfg.values()])
        n, cells = sizing(program, params or {})
        req = RunRequest(program.encode(), {"f64": 0, "f32": 1, "f32c": 2}[precision], len(names),
                         arr_n, arr_v, int(timesteps), (C.c_uint64 * 2)(*rng_state), n, cells,
                         len(cfg), arr_k, arr_c)
        rid = C.c_uint64()
        _check(self.lib.abl_server_submit(self.h, C.byref(req), C.byref(rid)))
        self._keep[rid.value] = (program, precision)
        return rid.value

    def status(self, rid: int) -> int:
        s = C.c_int()
        _check(self.lib.abl_server_status(self.h, rid, C.byref(s)))
        return s.value

    def wait(self, rid: int):
        _check(self.lib.abl_server_wait(self.h, rid))

    def result(self, rid: int, t: int = 0) -> np.ndarray:
        program, precision = self._keep[rid]
        dt = record_dtype(AGENTS[program][t][1], precision)
        n = C.c_size_t()
        _check(self.lib.abl_server_result(self.h, rid, t, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=dt)
        _check(self.lib.abl_server_result(self.h, rid, t, out.ctypes.data, n.value, C.byref(n)))
        return out[: n.value]

    def log(self, rid: int) -> str:
        n = C.c_size_t()
        _check(self.lib.abl_server_run_log(self.h, rid, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        _check(self.lib.abl_server_run_log(self.h, rid, buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    def times(self, rid: int):
        out = (C.c_double * 3)()
        _check(self.lib.abl_server_run_times(self.h, rid, out))
        return list(out)

    def stats(self) -> dict:
        out = (C.c_double * 6)()
        _check(self.lib.abl_server_stats(self.h, out))
        return {"completed": int(out[0]), "p50_s": out[1], "p99_s": out[2],
                "agent_steps": out[3], "span_s": out[4], "peak_bytes": int(out[5])}

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.abl_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
