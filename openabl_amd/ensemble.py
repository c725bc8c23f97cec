"""Python binding of batched ensembles (abl_ens_*, include/abl_hip.h).

Many independent runs of one program -- uneven agent counts, their own
parameters -- advanced together: one tiled step launch per timestep for all
members (through neighbour lists, re-sorted and rebuilt per member on the
timesteps a run of it alone would), replayed as a captured HIP graph.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .runtime import AGENTS, PREC, AblError, derived_params, load_library, record_dtype


def _check(rc: int):
    if rc != 0:
        raise AblError(rc, load_library().abl_ens_last_error().decode())


def member_seed(s0: int, s1: int, member: int):
    out = (C.c_uint64 * 2)()
    _check(load_library().abl_ens_member_seed(s0, s1, member, out))
    return int(out[0]), int(out[1])


class Ensemble:
    def __init__(self, program: str, precision: str, members: list[dict], config: dict | None = None):
        """members: one parameter dict per member (with "num_agents")."""
        self.lib = load_library()
        self.program = program
        self.precision = precision
        rows = [derived_params(program, m) for m in members]
        names = sorted({k for r in rows for k in r})
        for r in rows:
            missing = [k for k in names if k not in r]
            if missing:
                raise ValueError(f"every member must set the same parameters; missing {missing}")
        counts = (C.c_int * len(rows))(*[int(r["num_agents"]) for r in rows])
        cn = (C.c_char_p * len(names))(*[k.encode() for k in names])
        vals = np.array([[float(r[k]) for k in names] for r in rows], dtype=np.float64)
        self._vals = np.ascontiguousarray(vals)
        h = C.c_void_p()
        _check(self.lib.abl_ens_create(program.encode(), PREC[precision], len(rows), counts,
                                       len(names), cn, self._vals.ctypes.data, C.byref(h)))
        self.h = h
        self.members = len(rows)
        for k, v in (config or {}).items():
            _check(self.lib.abl_ens_set_config(self.h, k.encode(), str(v).encode()))

    def init(self, s0: int = 0xDEADBEEF, s1: int = 0xBEEFDEAD):
        _check(self.lib.abl_ens_init(self.h, s0, s1))

    def simulate(self, steps: int):
        _check(self.lib.abl_ens_simulate(self.h, int(steps)))

    def synchronize(self) -> float:
        ms = C.c_double()
        _check(self.lib.abl_ens_synchronize(self.h, C.byref(ms)))
        return ms.value

    def agents(self) -> int:
        n = C.c_size_t()
        _check(self.lib.abl_ens_agents(self.h, C.byref(n)))
        return n.value

    def download(self, member: int) -> np.ndarray:
        dt = record_dtype(AGENTS[self.program][0][1], self.precision)
        n = C.c_size_t()
        _check(self.lib.abl_ens_member_count(self.h, member, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=dt)
        got = C.c_size_t()
        _check(self.lib.abl_ens_download(self.h, member, out.ctypes.data, n.value, C.byref(got)))
        return out[: got.value]

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.abl_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
