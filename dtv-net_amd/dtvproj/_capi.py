"""ctypes binding of ``libdtvproj.so`` (declared in ``include/dtvproj.h``).

The library is loaded lazily on first use.  There is no fallback: if the
library is missing, was built from other sources than the ones beside it, or
a GPU entry point fails, a RuntimeError is raised.
"""
from __future__ import annotations

import ctypes
import os
import threading

from .build import LIB_PATH, source_digest

_lock = threading.Lock()
_lib = None

c_int32 = ctypes.c_int32
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_float_p = ctypes.POINTER(ctypes.c_float)


class DtvScheme(ctypes.Structure):
    _fields_ = [
        ("axis_mode", c_int32),
        ("reserved", c_int32),
        ("box_lo", c_double * 3),
        ("box_hi", c_double * 3),
        ("margin", c_double),
        ("scale", c_double),
    ]


class DtvLaunch(ctypes.Structure):
    _fields_ = [
        ("algorithm", c_int32),
        ("brick", c_int32 * 3),
        ("block", c_int32),
        ("view_parts", c_int32),
        ("reserved", c_int32 * 2),
    ]


ALGO_DEFAULT, ALGO_RAY, ALGO_BRICK = 0, 1, 2


def launch_struct(algorithm=0, brick=(0, 0, 0), block=0, view_parts=0) -> DtvLaunch:
    l = DtvLaunch()
    l.algorithm = int(algorithm)
    for i in range(3):
        l.brick[i] = int(brick[i])
    l.block = int(block)
    l.view_parts = int(view_parts)
    return l


# name -> argtypes; every entry point returns int status
_SIGS = {
    "dtv_standard_forward": [c_void_p, c_void_p, c_int32, c_void_p, c_int32, c_int32_p, c_int32_p,
                             c_double, c_double, c_int32, c_void_p],
    "dtv_standard_backward": [c_void_p, c_void_p, c_int32, c_void_p, c_int32, c_int32_p, c_int32_p,
                              c_double, c_double, c_int32, c_void_p],
    "dtv_standard_adjoint": [c_void_p, c_void_p, c_int32, c_void_p, c_int32, c_int32_p, c_int32_p,
                             c_double, c_double, c_int32, c_void_p],
    "dtv_beijing_forward": [c_void_p, c_void_p, c_int32, c_int32_p, c_int32_p, c_void_p, c_int32,
                            c_int32, c_void_p],
    "dtv_beijing_backward": [c_void_p, c_void_p, c_int32, c_int32_p, c_int32_p, c_void_p, c_int32,
                             c_int32, c_void_p],
    "dtv_beijing_adjoint": [c_void_p, c_void_p, c_int32, c_int32_p, c_int32_p, c_void_p, c_int32,
                            c_int32, c_void_p],
    "dtv_cone_forward": [c_void_p, c_void_p, c_int32, c_int32_p, c_int32_p, c_void_p, c_int32,
                         ctypes.POINTER(DtvScheme), c_void_p],
    "dtv_cone_adjoint": [c_void_p, c_void_p, c_int32, c_int32_p, c_int32_p, c_void_p, c_int32,
                         ctypes.POINTER(DtvScheme), c_void_p],
    "dtv_cone_forward_ex": [c_void_p, c_void_p, c_int32, c_int32_p, c_int32_p, c_void_p, c_int32,
                            ctypes.POINTER(DtvScheme), c_void_p, c_void_p],
    "dtv_cone_adjoint_ex": [c_void_p, c_void_p, c_int32, c_int32_p, c_int32_p, c_void_p, c_int32,
                            ctypes.POINTER(DtvScheme), c_void_p, c_void_p],
    "dtv_cone_forward_multi": [c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_int32,
                               c_void_p, c_int32, c_void_p],
    "dtv_cone_adjoint_multi": [c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_int32,
                               c_void_p, c_int32, c_void_p],
    "dtv_cone_adjoint_multi_ws": [c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_int32,
                                  c_int32, c_void_p, c_int32, c_void_p, ctypes.c_size_t,
                                  c_void_p],
    "dtv_cone_adjoint_multi_ws_bytes": [c_void_p, c_int32, ctypes.POINTER(ctypes.c_size_t)],
    "dtv_tv_fista_step": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int32_p,
                          ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                          c_int32, c_int32, c_void_p],
    "dtv_tv_fista_step_multi": [c_void_p, c_void_p, c_int32, c_int32, c_void_p],
    "dtv_cone_forward_ws": [c_void_p, c_void_p, c_int32, c_int32_p, c_int32_p, c_void_p, c_int32,
                            ctypes.POINTER(DtvScheme), c_void_p, c_void_p, ctypes.c_size_t,
                            c_void_p],
    "dtv_cone_forward_plan": [c_int32, c_int32_p, c_int32_p, c_int32, c_void_p,
                              ctypes.POINTER(DtvLaunch), ctypes.POINTER(ctypes.c_size_t)],
    "dtv_cone_adjoint_ws": [c_void_p, c_void_p, c_int32, c_int32_p, c_int32_p, c_void_p, c_int32,
                            ctypes.POINTER(DtvScheme), c_void_p, c_void_p, ctypes.c_size_t,
                            c_void_p],
    "dtv_cone_adjoint_plan": [c_int32, c_int32_p, c_int32_p, c_int32, c_void_p,
                              ctypes.POINTER(DtvLaunch), ctypes.POINTER(ctypes.c_size_t)],
    "dtv_tv_descent": [c_void_p, c_void_p, c_int32, c_int32_p, ctypes.c_float, ctypes.c_float,
                       c_int32, c_void_p],
    "dtv_pdhg_dtv_step": [c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int32_p,
                          ctypes.c_float, ctypes.c_float, c_float_p, c_int32, c_void_p],
    "dtv_l1ball": [c_void_p, c_int32, c_int32, c_double, c_int32, c_void_p, ctypes.c_size_t,
                   c_void_p],
    "dtv_l1ball_noalloc": [c_void_p, c_int32, c_int32, c_double, c_int32, c_void_p],
    "dtv_shrink_step": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int64,
                        ctypes.c_float, ctypes.c_float, ctypes.c_float, c_void_p],
    "dtv_fdpg_step": [c_void_p, c_void_p, c_void_p, c_int32, c_int32_p, ctypes.c_float,
                      c_float_p, c_float_p, ctypes.c_float, c_int32, c_void_p, ctypes.c_size_t,
                      c_void_p],
    "dtv_row_filter": [c_void_p, c_void_p, ctypes.c_int64, c_int32, c_void_p, c_int32, c_void_p,
                       c_int32, c_void_p],
    "dtv_views_standard": [c_double_p, c_int32, c_int32_p, c_int32_p, c_double, c_double,
                           c_double_p, ctypes.POINTER(DtvScheme)],
    "dtv_views_beijing": [c_float_p, c_int32, c_int32_p, c_int32_p, c_double_p,
                          ctypes.POINTER(DtvScheme)],
}

# entry points returning a size (no GPU work)
_SIZE_FNS = {"dtv_l1ball_workspace_bytes": [c_int32], "dtv_fdpg_workspace_bytes": []}

EXPORTED = tuple(_SIGS) + tuple(_SIZE_FNS) + ("dtv_strerror", "dtv_version")


def lib():
    """The loaded library (raises RuntimeError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"dtvproj HIP library not found at {LIB_PATH}; build it with "
                "`python -m dtvproj.build` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, argtypes in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = argtypes
            fn.restype = c_int32
        for name, argtypes in _SIZE_FNS.items():
            fn = getattr(L, name)
            fn.argtypes = argtypes
            fn.restype = ctypes.c_size_t
        L.dtv_strerror.argtypes = [c_int32]
        L.dtv_strerror.restype = ctypes.c_char_p
        L.dtv_version.argtypes = []
        L.dtv_version.restype = ctypes.c_char_p
        check_fresh(L.dtv_version().decode())
        _lib = L
        return _lib


def check_fresh(version_string: str) -> None:
    """Raise if the library's compiled-in source digest differs from the
    digest of the sources beside it (an edit since the last build)."""
    want = source_digest()
    if want is None:  # installed without its sources: nothing to compare
        return
    have = version_string.rpartition("src:")[2] if "src:" in version_string else None
    if have != want:
        raise RuntimeError(
            f"{LIB_PATH} was built from other sources (digest {have}, sources {want}); "
            "rebuild it with `python -m dtvproj.build`")


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().dtv_strerror(rc).decode()
        raise RuntimeError(f"{what} failed: {msg} (code {rc})")


def int3(vals):
    return (c_int32 * 3)(*[int(v) for v in vals])


def int2(vals):
    return (c_int32 * 2)(*[int(v) for v in vals])


def scheme_struct(scheme) -> DtvScheme:
    s = DtvScheme()
    s.axis_mode = int(scheme.axis_mode)
    s.reserved = 0
    for i in range(3):
        s.box_lo[i] = float(scheme.box_lo[i])
        s.box_hi[i] = float(scheme.box_hi[i])
    s.margin = float(scheme.margin)
    s.scale = float(scheme.scale)
    return s


def version() -> str:
    return lib().dtv_version().decode()
