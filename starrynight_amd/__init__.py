"""MI355X-native Metropolis dipole-lattice path of StarryNight.

The compute path is libstarrynight_amd.so (hand-written HIP kernels for gfx950
behind the C ABI in include/starrynight_amd.h); this package binds it with
ctypes.  There is no CPU fallback: without the built library every entry point
raises ImportError.
"""
from ._lib import EXPORTS, LIB_PATH, SnError, lib  # noqa: F401
from .lattice import (ReplicaParams, StarryNight, TuningParams, autotune, gen_neighbour,  # noqa: F401
                      plan_schedule, reference_initial_lattice)

__all__ = ["StarryNight", "ReplicaParams", "TuningParams", "SnError", "gen_neighbour", "plan_schedule",
           "reference_initial_lattice", "autotune", "lib",
           "LIB_PATH", "EXPORTS"]
