"""ctypes binding of libwavehip.so (the C ABI declared in include/wave_abi.h).

The library is built in-tree by ``make -C wave-simulation_amd`` (or
``__graft_entry__.build()``).  Loading fails loudly if it is missing: there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwavehip.so")

WV_OPS_SPARSE = 0
WV_OPS_STENCIL = 1
WV_PATH_FUSED = 0
WV_PATH_PLAIN = 1
WV_PATH_PHASED = 2
WV_PATH_STRIP = 3
WV_PATH_PERSISTENT = 4
WV_ENONFINITE = -5

# every entry point of include/wave_abi.h
EXPORTS = (
    "wv_create", "wv_destroy", "wv_init", "wv_set_model", "wv_set_acquisition", "wv_reset", "wv_run",
    "wv_synchronize", "wv_get_seismograms", "wv_get_wavefield", "wv_set_wavefield", "wv_last_run_ms",
    "wv_tune", "wv_set_kernel_timing", "wv_kernel_stats", "wv_last_error", "wv_set_checkpointing",
    "wv_set_adjoint_source", "wv_run_adjoint", "wv_get_gradient", "wv_set_adjoint_store",
    "wv_check_parameters", "wv_check_finite", "wv_check_finite_noalloc", "wv_adjoint_plan", "wv_source_signal", "wv_set_relaxation",
)


class Config(C.Structure):
    """wv_config (include/wave_abi.h)."""
    _fields_ = [
        ("NX", C.c_int), ("NY", C.c_int), ("NZ", C.c_int), ("NT", C.c_int),
        ("DH", C.c_double), ("DT", C.c_double),
        ("spatial_fd_order", C.c_int), ("ops_mode", C.c_int), ("free_surface", C.c_int),
        ("damping_boundary", C.c_int), ("boundary_width", C.c_int),
        ("damping_coeff", C.c_double), ("npower", C.c_double), ("fc_cpml", C.c_double), ("vmax_cpml", C.c_double),
        ("nshots", C.c_int), ("path", C.c_int), ("steps_per_graph", C.c_int), ("tile_y", C.c_int),
    ]


class WaveError(RuntimeError):
    """Raised for a non-zero status; carries the library's message (the reference throws)."""


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run `make -C wave-simulation_amd` or __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.wv_last_error.restype = C.c_char_p
        L.wv_create.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]
        L.wv_destroy.argtypes = [C.c_void_p]
        L.wv_init.argtypes = [C.c_void_p, C.POINTER(Config)]
        fp = C.POINTER(C.c_float)
        llp = C.POINTER(C.c_longlong)
        ip = C.POINTER(C.c_int)
        L.wv_set_model.argtypes = [C.c_void_p, fp, fp, fp]
        L.wv_set_acquisition.argtypes = [C.c_void_p, C.c_int, C.c_int, llp, ip, fp, C.c_int, llp, ip]
        L.wv_reset.argtypes = [C.c_void_p]
        L.wv_run.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.wv_synchronize.argtypes = [C.c_void_p]
        L.wv_get_seismograms.argtypes = [C.c_void_p, C.c_int, fp]
        L.wv_get_wavefield.argtypes = [C.c_void_p, C.c_int, C.c_int, fp]
        L.wv_set_wavefield.argtypes = [C.c_void_p, C.c_int, C.c_int, fp]
        L.wv_last_run_ms.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.wv_set_kernel_timing.argtypes = [C.c_void_p, C.c_int]
        L.wv_tune.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        L.wv_kernel_stats.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_double),
                                      C.POINTER(C.c_longlong), C.POINTER(C.c_double)]
        L.wv_set_checkpointing.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.wv_set_adjoint_source.argtypes = [C.c_void_p, C.c_int, fp]
        L.wv_set_adjoint_store.argtypes = [C.c_void_p, C.c_double]
        L.wv_check_finite_noalloc.argtypes = [C.c_void_p, llp]
        L.wv_adjoint_plan.argtypes = [C.c_void_p, ip, ip, ip]
        L.wv_set_relaxation.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), fp, fp]
        L.wv_source_signal.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, fp]
        L.wv_run_adjoint.argtypes = [C.c_void_p]
        L.wv_get_gradient.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
        L.wv_check_parameters.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(Config), fp, fp, fp, C.c_double, ip]
        L.wv_check_finite.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise WaveError(f"status {rc}: {load().wv_last_error().decode()}")
    return rc
