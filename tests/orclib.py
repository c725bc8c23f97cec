"""ctypes access to the oracle (test infrastructure only).

The oracle is the C restatement of the reference C backend's generated
programs (``oracle/models/*.c``) linked against the oracle's own restatement
of ``asset/c/libabl.{h,c}`` (``oracle/abl_rt.c``).  It runs the reference's
brute-force near loop (``src/backend/CPrinter.cpp:148-173``) or, switched at
run time, a grid that visits the same candidates in the same index order and
is therefore bit-identical (``oracle/orc_near.h``).

Every call loads a private copy of the model library, so instances do not
share the C globals (agent arrays, parameters, generator state).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import shutil
import subprocess
import tempfile

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
BUILD_DIR = os.path.join(ORACLE_DIR, "_build")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")

TYPE_END, TYPE_BOOL, TYPE_INT, TYPE_FLOAT, TYPE_STRING, TYPE_FLOAT2, TYPE_FLOAT3 = range(7)


class TypeInfo(C.Structure):
    _fields_ = [("type", C.c_int), ("offset", C.c_uint), ("name", C.c_char_p), ("is_pos", C.c_bool)]


def ensure_built() -> None:
    """Build the oracle libraries if they are missing (CPU only, gcc)."""
    need = os.path.join(BUILD_DIR, "liborc_circle_f64.so")
    if not os.path.exists(need):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR, "-j8"], check=True)


def fold6(x: float) -> float:
    """A folded float constant as the reference prints it: ostream default
    precision, i.e. 6 significant digits (src/backend/GenericPrinter.cpp:135-146)."""
    return float("%.6g" % x)


def member_dtype(info_ptr, float_bytes: int) -> np.dtype:
    """numpy dtype equal to the generated C struct, read from its RTTI."""
    fcode = "<f8" if float_bytes == 8 else "<f4"
    names, formats, offsets = [], [], []
    i = 0
    while True:
        m = info_ptr[i]
        if m.type == TYPE_END:
            itemsize = m.offset
            break
        names.append(m.name.decode())
        offsets.append(m.offset)
        formats.append({TYPE_BOOL: "u1", TYPE_INT: "<i4", TYPE_FLOAT: fcode,
                        TYPE_FLOAT2: (fcode, 2), TYPE_FLOAT3: (fcode, 3)}[m.type])
        i += 1
    return np.dtype({"names": names, "formats": formats, "offsets": offsets,
                     "itemsize": itemsize})


class Oracle:
    """One private instance of a restated model (precision 'f64' or 'f32')."""

    def __init__(self, model: str, precision: str = "f64", grid: bool = True):
        ensure_built()
        src = os.path.join(BUILD_DIR, f"liborc_{model}_{precision}.so")
        self._tmpdir = tempfile.mkdtemp(prefix="orc_")
        path = os.path.join(self._tmpdir, os.path.basename(src))
        shutil.copyfile(src, path)
        lib = C.CDLL(path)
        self.lib = lib
        self.model = model
        self.precision = precision
        lib.orc_set_param.argtypes = [C.c_char_p, C.c_double]
        lib.orc_get_param.argtypes = [C.c_char_p]
        lib.orc_get_param.restype = C.c_double
        lib.orc_agent_size.restype = C.c_size_t
        lib.orc_agent_count.restype = C.c_size_t
        lib.orc_agent_data.restype = C.c_void_p
        lib.orc_agent_set.argtypes = [C.c_int, C.c_void_p, C.c_size_t]
        lib.orc_agent_name.restype = C.c_char_p
        lib.orc_step_name.restype = C.c_char_p
        lib.orc_agent_info.restype = C.POINTER(TypeInfo)
        lib.orc_neighbours.argtypes = [C.c_int, C.c_void_p, C.c_long]
        lib.orc_neighbours.restype = C.c_long
        lib.orc_save.argtypes = [C.c_char_p]
        lib.orc_rng_set.argtypes = [C.c_uint64, C.c_uint64]
        lib.orc_rng_get.argtypes = [C.c_void_p]
        lib.orc_step_radius.restype = C.c_double
        lib.orc_env.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        self.float_bytes = lib.orc_float_bytes()
        self.ntypes = lib.orc_num_agent_types()
        self.type_names = [lib.orc_agent_name(t).decode() for t in range(self.ntypes)]
        self.dtypes = [member_dtype(lib.orc_agent_info(t), self.float_bytes) for t in range(self.ntypes)]
        for t in range(self.ntypes):
            assert self.dtypes[t].itemsize == lib.orc_agent_size(t)
        self.steps = [lib.orc_step_name(s).decode() for s in range(lib.orc_num_steps())]
        lib.orc_set_grid(1 if grid else 0)

    def close(self):
        shutil.rmtree(self._tmpdir, ignore_errors=True)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # parameters -----------------------------------------------------------
    def set(self, **params):
        for k, v in params.items():
            if self.lib.orc_set_param(k.encode(), float(v)) != 0:
                raise KeyError(k)
        return self

    def get(self, name: str) -> float:
        return self.lib.orc_get_param(name.encode())

    def set_grid(self, on: bool):
        self.lib.orc_set_grid(1 if on else 0)

    def rng_state(self):
        out = (C.c_uint64 * 2)()
        self.lib.orc_rng_get(out)
        return int(out[0]), int(out[1])

    def set_rng_state(self, s0: int, s1: int):
        self.lib.orc_rng_set(s0, s1)

    def env(self):
        lo = (C.c_double * 3)()
        hi = (C.c_double * 3)()
        g = C.c_double()
        self.lib.orc_env(lo, hi, C.byref(g))
        return list(lo), list(hi), g.value

    # state ----------------------------------------------------------------
    def init(self):
        self.lib.orc_init()
        return self

    def agents(self, t: int = 0) -> np.ndarray:
        n = self.lib.orc_agent_count(t)
        dt = self.dtypes[t]
        if n == 0:
            return np.zeros(0, dtype=dt)
        buf = (C.c_char * (n * dt.itemsize)).from_address(self.lib.orc_agent_data(t))
        return np.frombuffer(bytes(buf), dtype=dt).copy()

    def set_agents(self, arr: np.ndarray, t: int = 0):
        arr = np.ascontiguousarray(arr, dtype=self.dtypes[t])
        self.lib.orc_agent_set(t, arr.ctypes.data, len(arr))

    def simulate(self, timesteps: int):
        self.lib.orc_simulate(int(timesteps))

    def run_step(self, s: int):
        self.lib.orc_run_step(int(s))

    def neighbours(self, s: int = 0):
        """(centre, neighbour) index pairs the step's near loop accepts, in
        the order the reference visits them."""
        total = self.lib.orc_neighbours(s, None, 0)
        out = np.zeros(2 * max(total, 1), dtype=np.int64)
        got = self.lib.orc_neighbours(s, out.ctypes.data, total)
        assert got == total
        pairs = out[: 2 * total].reshape(-1, 2)
        return pairs[:, 0].copy(), pairs[:, 1].copy()

    def save(self, path: str):
        self.lib.orc_save(path.encode())

    # -b hip semantics of dynamic models (oracle/orc_dynamic.c) ------------
    def ids(self, t: int = 0) -> np.ndarray:
        self.lib.orc_agent_ids.restype = C.POINTER(C.c_longlong)
        n = self.lib.orc_agent_count(t)
        if n == 0:
            return np.zeros(0, dtype=np.int64)
        return np.ctypeslib.as_array(self.lib.orc_agent_ids(t), shape=(n,)).astype(np.int64)

    def log_text(self) -> str:
        self.lib.orc_log_text.restype = C.c_char_p
        return self.lib.orc_log_text().decode()

    def set_seed(self, seed: int):
        self.lib.orc_set_seed.argtypes = [C.c_ulonglong]
        self.lib.orc_set_seed(seed)


def ref_binary(model: str, precision: str = "f64") -> str | None:
    """Path of the reference-runtime build of a restated model, if built."""
    p = os.path.join(REF_DIR, f"{model}_{precision}")
    return p if os.path.exists(p) else None


# Parameters of the benchmark configurations, with derived constants folded
# the way the reference compiler folds them (src/AnalysisVisitor.cpp:403-472).
def circle_params(num_agents=1024, rho=0.1024, r=2.5, k_rep=0.05, k_att=0.01, num_timesteps=100):
    return dict(num_agents=num_agents, rho=rho, r=r, k_rep=k_rep, k_att=k_att,
                num_timesteps=num_timesteps, W=fold6(math.sqrt(num_agents / rho)))


def circle3d_params(num_agents=1000, rho=0.05, r=5.0, num_clusters=0, cluster_sigma=20.0,
                    background_fraction=0.2, k_rep=0.05, k_att=0.01, num_timesteps=100):
    return dict(num_agents=num_agents, rho=rho, r=r, k_rep=k_rep, k_att=k_att,
                num_clusters=num_clusters, cluster_sigma=cluster_sigma,
                background_fraction=background_fraction, num_timesteps=num_timesteps,
                W=fold6(float(np.cbrt(num_agents / rho))))


def boids2d_params(num_agents=65536, agent_density=0.065536, interaction_radius=10.0,
                   separation_radius=2.0, steer_scale=1.0, collision_scale=1.5,
                   match_scale=1.0, max_speed=2.0, time_scale=0.0005, global_scale=0.15,
                   num_timesteps=200):
    return dict(num_agents=num_agents, agent_density=agent_density,
                interaction_radius=interaction_radius, separation_radius=separation_radius,
                steer_scale=steer_scale, collision_scale=collision_scale,
                match_scale=match_scale, max_speed=max_speed, time_scale=time_scale,
                global_scale=global_scale, num_timesteps=num_timesteps,
                min_pos=0.0, max_pos=fold6(math.sqrt(num_agents / agent_density)))


def predator_prey_params(num_agents=3200, prey_fraction=0.35, predator_fraction=0.025,
                         agent_density=1600.0, SPACE_MULT=1.0, num_timesteps=100, **extra):
    p = dict(num_agents=num_agents, prey_fraction=prey_fraction,
             predator_fraction=predator_fraction, agent_density=agent_density,
             SPACE_MULT=SPACE_MULT, num_timesteps=num_timesteps,
             PRED_PREY_INTERACTION_RADIUS=fold6(0.1 * SPACE_MULT),
             PREY_GROUP_COHESION_RADIUS=fold6(0.12 * SPACE_MULT),
             SAME_SPECIES_AVOIDANCE_RADIUS=fold6(0.035 * SPACE_MULT),
             GRASS_EAT_DISTANCE=fold6(0.02 * SPACE_MULT),
             PRED_KILL_DISTANCE=fold6(0.02 * SPACE_MULT),
             env_size=fold6(math.sqrt(num_agents / agent_density) * SPACE_MULT))
    p.update(extra)
    return p


def game_of_life_params(num_agents=400, alive_fraction=0.1, debug=False, num_timesteps=100):
    return dict(num_agents=num_agents, alive_fraction=alive_fraction, debug=int(debug),
                num_timesteps=num_timesteps, size=int(math.sqrt(num_agents)))


def field_bytes(arr: np.ndarray) -> bytes:
    """Bytes of every named field (struct padding excluded)."""
    return b"".join(np.ascontiguousarray(arr[name]).tobytes() for name in arr.dtype.names)
