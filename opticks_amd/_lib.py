"""
ctypes binding of the C-ABI in include/opticks_amd.h (libopticks_amd.so).

The product path is the HIP library: importing this module when the library
has not been built raises ImportError (no CPU fallback exists).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# OPTICKS_AMD_LIB selects another build of the same library (the OA_STATS counters
# build used by opticks_amd.workcount); the default is the in-tree libopticks_amd.so
LIBPATH = os.environ.get("OPTICKS_AMD_LIB") or os.path.join(HERE, "libopticks_amd.so")

MI_OK, MI_NOTHING_TO_DO = 0, 1
MI_OUT_PHOTON, MI_OUT_SEQ, MI_OUT_HIT = 0x1, 0x2, 0x4
MI_RNG_PHILOX, MI_RNG_PRECOOKED, MI_RNG_XORWOW = 0, 1, 2
MI_KERNEL_PLAIN, MI_KERNEL_WAVEFRONT = 0, 1


class MiGeometry(C.Structure):
    _fields_ = [("node", C.c_void_p), ("n_node", C.c_int64),
                ("prim", C.c_void_p), ("n_prim", C.c_int64),
                ("solid", C.c_void_p), ("n_solid", C.c_int64),
                ("itra", C.c_void_p), ("n_tran", C.c_int64),
                ("plan", C.c_void_p), ("n_plan", C.c_int64),
                ("inst", C.c_void_p), ("n_inst", C.c_int64)]


class MiPhysics(C.Structure):
    _fields_ = [("bnd", C.c_void_p), ("n_bnd", C.c_int32), ("n_wl", C.c_int32),
                ("wl_low", C.c_float), ("wl_step", C.c_float), ("optical", C.c_void_p),
                ("icdf", C.c_void_p), ("n_icdf_layer", C.c_int32), ("n_icdf", C.c_int32),
                ("icdf_hd_factor", C.c_int32)]


class MiConfig(C.Structure):
    _fields_ = [("max_bounce", C.c_int32), ("max_time", C.c_float), ("tmin", C.c_float),
                ("tmin0", C.c_float), ("eps0_mask", C.c_uint32), ("tmax", C.c_float),
                ("hit_mask", C.c_uint32), ("seed", C.c_uint64), ("rng_offset", C.c_uint64),
                ("event_skipahead", C.c_uint64), ("rng_kind", C.c_int32),
                ("precooked", C.c_void_p), ("precooked_ni", C.c_int64), ("precooked_nk", C.c_int64),
                ("output_mask", C.c_uint32), ("max_slot", C.c_int64), ("kernel", C.c_int32),
                ("block_size", C.c_int32), ("refill_period", C.c_int32), ("bvh_leaf_size", C.c_int32),
                ("waves_per_cu", C.c_int32), ("trace_regroup_pct", C.c_int32), ("refill_batch", C.c_int32),
                ("adjoint_tape", C.c_int32)]


class MiEventOut(C.Structure):
    _fields_ = [("n_photon", C.c_int64), ("n_hit", C.c_int64), ("photon", C.c_void_p),
                ("seq", C.c_void_p), ("hit", C.c_void_p), ("device_ms", C.c_double)]


class MiAdjoint(C.Structure):
    _fields_ = [("mat_block", C.c_void_p), ("n_mat", C.c_int32), ("sample_bin", C.c_void_p),
                ("n_bin", C.c_int32), ("n_block", C.c_int32), ("n_det", C.c_int32),
                ("max_row_floats", C.c_int64)]


class MiTuneResult(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("refill_period", C.c_int32), ("bvh_leaf_size", C.c_int32),
                ("kernel", C.c_int32), ("trace_regroup_pct", C.c_int32), ("refill_batch", C.c_int32),
                ("ms", C.c_double)]


MI_RECORD_PHOTON, MI_RECORD_PHOTONLITE = 0, 1
MI_MERGE_ALREADY_SELECTED = 0xFFFFFFFF

SYMBOLS = ["mi_config_default", "mi_last_error", "mi_version", "mi_sim_create", "mi_sim_destroy",
           "mi_sim_set_config", "mi_sim_get_config", "mi_sim_simulate", "mi_sim_reset",
           "mi_sim_simulate_device", "mi_sim_simulate_batch_device", "mi_sim_trace_device",
           "mi_sim_refit_instances", "mi_sim_last_timing", "mi_sim_set_bnd", "mi_sim_adjoint_device",
           "mi_sim_last_adjoint_timing", "mi_batch_create", "mi_batch_launch", "mi_batch_wait",
           "mi_batch_num_photon", "mi_batch_destroy", "mi_sim_autotune", "mi_batch_create_adjoint",
           "mi_batch_last_timing", "mi_sim_merge_hits_device", "mi_sim_simulate_lite_device",
           "mi_sim_simtrace_device", "mi_sim_simtrace", "mi_sim_simulate_device_noalloc"]

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError(f"{LIBPATH} is missing: run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    L = C.CDLL(LIBPATH)
    L.mi_config_default.argtypes = [C.POINTER(MiConfig)]
    L.mi_last_error.restype = C.c_char_p
    L.mi_version.restype = C.c_char_p
    L.mi_sim_create.argtypes = [C.POINTER(MiGeometry), C.POINTER(MiPhysics), C.POINTER(MiConfig),
                                C.POINTER(C.c_void_p)]
    L.mi_sim_destroy.argtypes = [C.c_void_p]
    L.mi_sim_set_config.argtypes = [C.c_void_p, C.POINTER(MiConfig)]
    L.mi_sim_get_config.argtypes = [C.c_void_p, C.POINTER(MiConfig)]
    L.mi_sim_simulate.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.POINTER(MiEventOut)]
    L.mi_sim_reset.argtypes = [C.c_void_p, C.c_int32]
    L.mi_sim_simulate_device.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_int64, C.c_int64,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                         C.POINTER(C.c_int64), C.c_void_p]
    L.mi_sim_simulate_device_noalloc.argtypes = L.mi_sim_simulate_device.argtypes
    L.mi_sim_simulate_batch_device.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                               C.c_void_p, C.c_void_p]
    L.mi_sim_trace_device.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
    L.mi_sim_simtrace_device.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p]
    L.mi_sim_simtrace.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.POINTER(C.c_float)),
                                  C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L.mi_sim_refit_instances.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    L.mi_sim_last_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.mi_sim_set_bnd.argtypes = [C.c_void_p, C.c_void_p]
    L.mi_sim_adjoint_device.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_int64, C.c_int64,
                                        C.c_void_p, C.c_int64, C.POINTER(MiAdjoint), C.c_void_p, C.c_void_p]
    L.mi_sim_last_adjoint_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.mi_batch_create.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(C.c_void_p)]
    L.mi_batch_launch.argtypes = [C.c_void_p, C.c_void_p]
    L.mi_batch_create_adjoint.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(MiAdjoint), C.c_void_p,
                                          C.c_void_p, C.POINTER(C.c_void_p)]
    L.mi_batch_last_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.mi_sim_simulate_lite_device.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_int64, C.c_int64,
                                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                              C.POINTER(C.c_int64), C.c_void_p]
    L.mi_sim_merge_hits_device.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_uint32, C.c_float,
                                           C.c_void_p, C.c_int64, C.POINTER(C.c_int64), C.c_void_p]
    L.mi_batch_wait.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]
    L.mi_batch_num_photon.argtypes = [C.c_void_p]
    L.mi_batch_num_photon.restype = C.c_int64
    L.mi_batch_destroy.argtypes = [C.c_void_p]
    L.mi_sim_autotune.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.POINTER(MiTuneResult),
                                  C.c_int32, C.POINTER(C.c_int32)]
    for name in SYMBOLS:
        if name not in ("mi_last_error", "mi_version", "mi_config_default", "mi_sim_destroy"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def last_error() -> str:
    return lib().mi_last_error().decode()


def check(rc, what):
    if rc < 0:
        raise RuntimeError(f"{what} failed rc={rc}: {last_error()}")
    return rc
