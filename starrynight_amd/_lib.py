"""ctypes binding of libstarrynight_amd.so (the C ABI in include/starrynight_amd.h).

The library is loaded from this package directory only.  There is no fallback:
if the .so is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libstarrynight_amd.so")

SN_OK = 0
STATUS = {0: "SN_OK", 1: "SN_ERR_INVALID", 2: "SN_ERR_HIP", 3: "SN_ERR_NOMEM", 4: "SN_ERR_RANGE",
          5: "SN_ERR_UNSUPPORTED"}

# every entry point declared in include/starrynight_amd.h
EXPORTS = [
    "sn_create", "sn_destroy", "sn_last_error", "sn_status_string", "sn_set_lattice",
    "sn_set_lattice_device", "sn_gradient_device", "sn_gradient_all_device", "sn_gradient_all_device_noalloc",
    "sn_get_lattice", "sn_lattice_device_ptr", "sn_init_lattice_random", "sn_set_params",
    "sn_site_energy_batch", "sn_site_energy_batch_device", "sn_mc_sweeps", "sn_sync",
    "sn_mc_replay", "sn_counters", "sn_reset_counters", "sn_get_sweep_index", "sn_set_sweep_index",
    "sn_energy", "sn_gradient", "sn_moment",
    "sn_radial_order", "sn_dipole_potential", "sn_dipole_efield", "sn_recombination",
    "sn_schedule_info", "sn_schedule_order", "sn_stream", "sn_sweep_config", "sn_kernel_choice",
    "sn_time_sweep_kernel", "sn_plan_schedule", "sn_autotune", "sn_reference_initial_lattice",
]


class SnError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Params(C.Structure):
    _fields_ = [("X", C.c_int32), ("Y", C.c_int32), ("Z", C.c_int32), ("dipole_cutoff", C.c_int32),
                ("cage_strain", C.c_double), ("efield", C.c_double * 3), ("K", C.c_double),
                ("constrain_to_x", C.c_int32), ("temperature", C.c_double), ("seed", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("tile", C.c_int32 * 3), ("lanes_per_site", C.c_int32),
                ("fine_colouring", C.c_int32), ("use_graph", C.c_int32), ("kernel", C.c_int32),
                ("block_threads", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run `python -m starrynight_amd.build` "
                "(the HIP library is required, there is no CPU fallback)")
        # torch ships a HIP runtime with the same soname as ROCm's: whichever
        # loads first serves both.  torch set up after this library's runtime
        # has reported no GPUs, so torch (when installed) loads first and the
        # two share its runtime, which also lets device pointers pass between
        # them (the *_device entry points).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        P, vp = C.POINTER, C.c_void_p
        dp, ip, u8p, u64p = P(C.c_double), P(C.c_int32), P(C.c_uint8), P(C.c_uint64)
        L.sn_create.argtypes = [P(Params), C.c_int32, P(Tuning), P(vp)]
        L.sn_destroy.argtypes = [vp]
        L.sn_last_error.argtypes = [vp]; L.sn_last_error.restype = C.c_char_p
        L.sn_status_string.argtypes = [C.c_int]; L.sn_status_string.restype = C.c_char_p
        L.sn_set_lattice.argtypes = [vp, C.c_int32, dp]
        L.sn_get_lattice.argtypes = [vp, C.c_int32, dp]
        L.sn_lattice_device_ptr.argtypes = [vp, C.c_int32, P(vp)]
        L.sn_init_lattice_random.argtypes = [vp, C.c_int32, C.c_int32, dp, dp]
        L.sn_set_params.argtypes = [vp, C.c_int32, C.c_double, dp, C.c_double]
        L.sn_site_energy_batch.argtypes = [vp, C.c_int32, C.c_int64, ip, dp, dp]
        L.sn_site_energy_batch_device.argtypes = [vp, C.c_int32, C.c_int64, vp, vp, vp]
        L.sn_mc_sweeps.argtypes = [vp, C.c_int32]
        L.sn_sync.argtypes = [vp]
        L.sn_mc_replay.argtypes = [vp, C.c_int32, C.c_int64, ip, dp, dp, u8p, u8p, u8p, u64p, u64p]
        L.sn_counters.argtypes = [vp, C.c_int32, u64p, u64p, u64p]
        L.sn_reset_counters.argtypes = [vp]
        L.sn_get_sweep_index.argtypes = [vp, C.c_int32, u64p]
        L.sn_set_sweep_index.argtypes = [vp, C.c_int32, C.c_uint64]
        L.sn_energy.argtypes = [vp, C.c_int32, dp, dp]
        L.sn_gradient.argtypes = [vp, C.c_int32, dp, dp, dp]
        L.sn_gradient_device.argtypes = [vp, C.c_int32, vp, dp, dp]
        L.sn_gradient_all_device.argtypes = [vp, vp, dp, dp]
        L.sn_gradient_all_device_noalloc.argtypes = [vp, vp, dp, dp]
        L.sn_set_lattice_device.argtypes = [vp, C.c_int32, vp]
        L.sn_moment.argtypes = [vp, C.c_int32, dp]
        L.sn_radial_order.argtypes = [vp, C.c_int32, dp, dp, C.POINTER(C.c_int64)]
        L.sn_dipole_potential.argtypes = [vp, C.c_int32, dp]
        L.sn_dipole_efield.argtypes = [vp, C.c_int32, C.c_int32, dp]
        L.sn_recombination.argtypes = [vp, C.c_int32, dp]
        L.sn_schedule_info.argtypes = [vp, C.c_int32, P(C.c_int64), ip]
        L.sn_schedule_order.argtypes = [vp, C.c_int32, ip, ip]
        L.sn_stream.argtypes = [vp, P(vp)]
        L.sn_sweep_config.argtypes = [vp, ip, ip, ip, ip]
        L.sn_kernel_choice.argtypes = [vp, ip, ip, ip, ip]
        L.sn_time_sweep_kernel.argtypes = [vp, C.c_int32, dp, ip]
        L.sn_autotune.argtypes = [P(Params), C.c_int32, C.c_int32, dp, dp, C.c_int32, C.c_int32, P(Tuning), dp]
        L.sn_reference_initial_lattice.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_char_p, C.c_uint32,
                                                   C.c_int32, dp, dp, dp]
        L.sn_plan_schedule.argtypes = [P(Params), C.c_int32, P(Tuning), C.c_int32, P(C.c_int64), ip,
                                       ip, ip]
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))
