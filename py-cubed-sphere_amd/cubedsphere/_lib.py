"""ctypes binding of libcsadv.so (the C ABI declared in include/csadv.h).

The shared library is built in-tree next to this file by ``__graft_entry__.build()``.
There is no fallback: if the library is missing every solver call raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcsadv.so")

CSADV_F64, CSADV_F32 = 0, 1
ABI_VERSION = 7          # CSADV_ABI_VERSION of include/csadv.h this binding mirrors
_ERRORS = {-1: "invalid argument", -2: "HIP runtime error", -3: "call out of order", -4: "size out of range"}

# exported symbols, in declaration order of include/csadv.h
SYMBOLS = (
    "csadv_create", "csadv_partials_per_tracer", "csadv_work_bytes", "csadv_tile_rows", "csadv_update", "csadv_wind_init",
    "csadv_ghost_fill", "csadv_step", "csadv_set_graph", "csadv_finalize", "csadv_mass", "csadv_norms",
    "csadv_diagnostics", "csadv_diagnostics_noalloc", "csadv_status",
    "csadv_current", "csadv_profile", "csadv_profile_read", "csadv_time_step_kernel", "csadv_time_step_kernel_rotating", "csadv_remap_latlon", "csadv_last_error", "csadv_destroy",
    "csadv_abi_version",
)

_p = ctypes.c_void_p


class SolveDesc(ctypes.Structure):
    """Mirror of ``csadv_solve_desc`` (include/csadv.h)."""

    _fields_ = [
        ("N", ctypes.c_int32), ("ntr", ctypes.c_int32), ("dtype", ctypes.c_int32),
        ("ld", ctypes.c_int32), ("ldu", ctypes.c_int32),
        ("recon", ctypes.c_int32), ("dp", ctypes.c_int32), ("split", ctypes.c_int32),
        ("et", ctypes.c_int32), ("mt", ctypes.c_int32), ("mf", ctypes.c_int32),
        ("nbasis", ctypes.c_int32), ("unsteady", ctypes.c_int32), ("max_steps", ctypes.c_int32),
        ("dt", ctypes.c_double), ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("a2", ctypes.c_double),
        ("q0", _p), ("q1", _p), ("uavg", _p), ("vavg", _p), ("usgn", _p), ("vsgn", _p),
        ("gc", _p), ("gu", _p), ("gv", _p), ("rgc", _p), ("gfix", _p), ("gc64", _p),
        ("lagw", _p), ("lagk", _p), ("orient", _p), ("wbu", _p), ("wbv", _p),
        ("conv_pc", _p), ("conv_pu", _p), ("conv_pv", _p), ("wpc", _p),
        ("n_strip", ctypes.c_int32), ("n_corner", ctypes.c_int32),
        ("strip_tgt", _p), ("strip_src", _p), ("strip_w", _p),
        ("corner_tgt", _p), ("corner_src", _p), ("corner_w", _p),
        ("coef", _p), ("corr", _p), ("partials", _p), ("ctl", _p),
        ("tile", ctypes.c_int32), ("stop", ctypes.c_int32), ("work", _p), ("offset", _p),
    ]


class CsadvError(RuntimeError):
    pass


_lib = None


def load():
    """Load libcsadv.so once; raise if it is absent (no silent fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CsadvError("libcsadv.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    vp, i32, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)
    lib.csadv_create.argtypes = [ctypes.POINTER(SolveDesc), i32, i32, ctypes.POINTER(vp)]
    lib.csadv_partials_per_tracer.argtypes = [i32, i32]
    lib.csadv_tile_rows.argtypes = [i32]
    lib.csadv_work_bytes.argtypes = [i32, i32, ctypes.POINTER(ctypes.c_int64)]
    lib.csadv_update.argtypes = [vp, ctypes.POINTER(SolveDesc), i32, vp]
    lib.csadv_wind_init.argtypes = [vp, vp]
    lib.csadv_ghost_fill.argtypes = [vp, vp]
    lib.csadv_step.argtypes = [vp, i32, vp]
    lib.csadv_set_graph.argtypes = [vp, i32]
    lib.csadv_finalize.argtypes = [vp, i32, vp]
    lib.csadv_mass.argtypes = [vp, i32, i32, dbl, vp]
    lib.csadv_norms.argtypes = [vp, i32, i32, vp, dbl, vp]
    lib.csadv_diagnostics.argtypes = [vp, i32, i32, i32, vp, ctypes.c_double, vp, vp]
    lib.csadv_diagnostics_noalloc.argtypes = [vp, i32, i32, i32, vp, ctypes.c_double, vp, vp]
    lib.csadv_status.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), vp]
    lib.csadv_current.argtypes = [vp, i32]
    lib.csadv_profile.argtypes = [vp, i32]
    lib.csadv_profile_read.argtypes = [vp, dbl, ctypes.POINTER(ctypes.c_int32)]
    lib.csadv_time_step_kernel.argtypes = [vp, i32, vp, dbl]
    lib.csadv_time_step_kernel_rotating.argtypes = [ctypes.POINTER(vp), i32, i32, vp, dbl]
    lib.csadv_remap_latlon.argtypes = [vp, i32, i32, vp, i32, i32, vp, vp, vp, vp, vp]
    lib.csadv_last_error.argtypes = [vp]
    lib.csadv_last_error.restype = ctypes.c_char_p
    lib.csadv_destroy.argtypes = [vp]
    lib.csadv_destroy.restype = None
    lib.csadv_abi_version.argtypes = []
    for name in SYMBOLS:
        if name not in ("csadv_last_error", "csadv_destroy"):
            getattr(lib, name).restype = ctypes.c_int
    if lib.csadv_abi_version() != ABI_VERSION:
        raise CsadvError("libcsadv ABI version mismatch")
    _lib = lib
    return lib


def check(rc, handle=None, what="csadv"):
    if rc >= 0:
        return rc
    msg = ""
    if handle:
        msg = (load().csadv_last_error(handle) or b"").decode(errors="replace")
    raise CsadvError("%s failed (%d: %s) %s" % (what, rc, _ERRORS.get(rc, "?"), msg))
