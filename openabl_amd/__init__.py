"""openabl_amd -- an MI355X (gfx950) backend for the OpenABL agent-based
modelling language.

The device runtime (``csrc/``, built to ``lib/libabl_hip.so``) holds agents
as cell-sorted struct-of-arrays in HBM and runs the per-timestep update of an
``.abl`` model -- neighbour loops, the simulate loop -- as hand-written HIP
kernels behind the C ABI of ``include/abl_hip.h``.  :mod:`openabl_amd.runtime`
binds that ABI for Python callers.
"""
from .runtime import (AGENT_NAMES, AGENTS, MEMBERS, AblError, Simulation, derived_params, fold6,
                      load_library, record_dtype)

__all__ = ["AGENT_NAMES", "AGENTS", "MEMBERS", "AblError", "Simulation", "derived_params",
           "fold6", "load_library", "record_dtype"]
