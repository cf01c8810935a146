"""ctypes binding of libeik's C ABI (include/eik.h).

The library is built in-tree (``eikonalsolvers.jl_amd/build_libeik.py``) next to this file.  There
is no CPU fallback: if ``libeik.so`` is missing or no GPU is visible, calls raise.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libeik.so")

EIK_OK = 0
EIK_EASSERT = -1
EIK_EINVAL = -2
EIK_ENOMEM = -3
EIK_EDEVICE = -4
EIK_ENOCONV = -5
EIK_ESIZE = -6

# every symbol include/eik.h declares
EXPORTS = (
    "eik_ctx_create", "eik_ctx_destroy", "eik_last_error", "eik_version",
    "eik_model_create", "eik_model_update_vel", "eik_model_destroy",
    "eik_traveltime", "eik_gradient", "eik_misfit", "eik_last_stats",
    "eik_server_create", "eik_server_destroy", "eik_submit", "eik_wait", "eik_server_stats",
    "eik_tracerays", "eik_traveltime_batch", "eik_poll", "eik_tracerays_max_steps", "eik_wait_stats",
    "eik_request_bytes",
    "eik_traveltime_noalloc", "eik_mem_info", "eik_ctx_set_streams",
)


class EikGrid(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("ndim", ctypes.c_int32), ("n", ctypes.c_int64 * 3),
                ("d", ctypes.c_double * 3), ("o", ctypes.c_double * 3)]


class EikParams(ctypes.Structure):
    _fields_ = [("refinearoundsrc", ctypes.c_int32), ("downscalefactor", ctypes.c_int32),
                ("noderadius", ctypes.c_int32), ("radiussmoothgradsrc", ctypes.c_int32),
                ("smoothgradkern", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("fast", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class EikRequest(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("whichgrad", ctypes.c_int32), ("model", ctypes.c_void_p),
                ("nsrc", ctypes.c_int64), ("coordsrc", ctypes.POINTER(ctypes.c_double)),
                ("rec_off", ctypes.POINTER(ctypes.c_int64)), ("coordrec", ctypes.POINTER(ctypes.c_double)),
                ("pickobs", ctypes.POINTER(ctypes.c_double)), ("stdobs", ctypes.POINTER(ctypes.c_double)),
                ("params", EikParams), ("ttpicks", ctypes.POINTER(ctypes.c_double)),
                ("gradvel", ctypes.POINTER(ctypes.c_double)), ("gradsrc", ctypes.POINTER(ctypes.c_double)),
                ("misf", ctypes.POINTER(ctypes.c_double))]


class LibEikMissing(RuntimeError):
    pass


_lib = None
_lock = threading.Lock()


def load():
    """Load libeik.so and declare the C signatures (no device call)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIBPATH):
            raise LibEikMissing(f"{LIBPATH} not found: build it with eikonalsolvers.jl_amd/build_libeik.py")
        L = ctypes.CDLL(LIBPATH)
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        dp = P(ctypes.c_double)
        ip = P(ctypes.c_int64)
        L.eik_ctx_create.argtypes = [ctypes.c_int32, P(vp)]
        L.eik_ctx_create.restype = ctypes.c_int
        L.eik_ctx_destroy.argtypes = [vp]
        L.eik_ctx_destroy.restype = None
        L.eik_last_error.argtypes = [vp]
        L.eik_last_error.restype = ctypes.c_char_p
        L.eik_version.argtypes = []
        L.eik_version.restype = ctypes.c_char_p
        L.eik_model_create.argtypes = [vp, P(EikGrid), dp, P(vp)]
        L.eik_model_create.restype = ctypes.c_int
        L.eik_model_update_vel.argtypes = [vp, dp]
        L.eik_model_update_vel.restype = ctypes.c_int
        L.eik_model_destroy.argtypes = [vp]
        L.eik_model_destroy.restype = None
        L.eik_traveltime.argtypes = [vp, vp, ctypes.c_int64, dp, ip, dp, P(EikParams), dp, dp]
        L.eik_traveltime.restype = ctypes.c_int
        L.eik_traveltime_noalloc.argtypes = [vp, vp, ctypes.c_int64, dp, ip, dp, P(EikParams), dp, dp]
        L.eik_traveltime_noalloc.restype = ctypes.c_int
        L.eik_mem_info.argtypes = [vp, ip, ip]
        L.eik_mem_info.restype = ctypes.c_int
        L.eik_ctx_set_streams.argtypes = [vp, ctypes.c_int32]
        L.eik_ctx_set_streams.restype = ctypes.c_int
        L.eik_gradient.argtypes = [vp, vp, ctypes.c_int64, dp, ip, dp, dp, dp, ctypes.c_int32, P(EikParams),
                                   dp, dp, dp]
        L.eik_gradient.restype = ctypes.c_int
        L.eik_misfit.argtypes = [vp, vp, ctypes.c_int64, dp, ip, dp, dp, dp, P(EikParams), dp]
        L.eik_misfit.restype = ctypes.c_int
        L.eik_last_stats.argtypes = [vp, dp, ctypes.c_int32]
        L.eik_last_stats.restype = ctypes.c_int
        L.eik_server_create.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, P(vp)]
        L.eik_server_create.restype = ctypes.c_int
        L.eik_server_destroy.argtypes = [vp]
        L.eik_server_destroy.restype = None
        L.eik_submit.argtypes = [vp, P(EikRequest), ip]
        L.eik_submit.restype = ctypes.c_int
        L.eik_wait.argtypes = [vp, ctypes.c_int64, dp, ctypes.c_char_p, ctypes.c_int64]
        L.eik_wait.restype = ctypes.c_int
        L.eik_wait_stats.argtypes = [vp, ctypes.c_int64, dp, dp, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int64]
        L.eik_wait_stats.restype = ctypes.c_int
        L.eik_request_bytes.argtypes = [P(EikRequest), ip]
        L.eik_request_bytes.restype = ctypes.c_int
        L.eik_server_stats.argtypes = [vp, ip, ctypes.c_int32]
        L.eik_server_stats.restype = ctypes.c_int
        L.eik_tracerays.argtypes = [vp, P(EikGrid), ctypes.c_int64, dp, dp, ip, dp, ctypes.c_double, ip, dp,
                                    ctypes.c_int64]
        L.eik_tracerays.restype = ctypes.c_int
        L.eik_traveltime_batch.argtypes = [vp, ctypes.c_int32, P(EikRequest)]
        L.eik_traveltime_batch.restype = ctypes.c_int
        L.eik_poll.argtypes = [vp, ctypes.c_int64, P(ctypes.c_int32)]
        L.eik_poll.restype = ctypes.c_int
        L.eik_tracerays_max_steps.argtypes = [P(EikGrid), ctypes.c_double, ip]
        L.eik_tracerays_max_steps.restype = ctypes.c_int
        _lib = L
        return L


def check(rc: int, ctx, msg: str | None = None) -> None:
    """Raise the exception kind the reference raises for the same failure."""
    if rc == EIK_OK:
        return
    if msg is None:
        msg = load().eik_last_error(ctx).decode(errors="replace") if ctx else "libeik error"
    if rc == EIK_EASSERT:
        raise AssertionError(msg)
    if rc == EIK_EINVAL:
        raise ValueError(msg)
    if rc == EIK_ENOMEM:
        raise MemoryError(msg)
    raise RuntimeError(f"libeik error {rc}: {msg}")


def ptr(a: np.ndarray, ct=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ct))
