"""ctypes binding of ``libinelastic_hip.so`` (the C ABI in include/inelastic.h).

The library is required: there is no CPU fallback.  Loading fails loudly with
the build command when the shared object is missing.
"""
import ctypes
import os
import pathlib

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "libinelastic_hip.so"

IB_OK = 0
IB_ERR_ARG = -1
IB_ERR_HIP = -2
IB_ERR_UNSUPPORTED = -3
IB_ERR_BC = -4
PREC_FP64 = 0
PREC_FP32 = 1
PREC_FP32_KAHAN = 2


class ib_plan_desc(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("n", ctypes.c_int),
        ("nr", ctypes.c_int),
        ("m", ctypes.c_int),
        ("L", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("n_param", ctypes.c_int),
        ("r", ctypes.POINTER(ctypes.c_double)),
        ("wr", ctypes.POINTER(ctypes.c_double)),
        ("sigma", ctypes.POINTER(ctypes.c_double)),
        ("wsigma", ctypes.POINTER(ctypes.c_double)),
        ("e_r", ctypes.POINTER(ctypes.c_double)),
    ]


_c = ctypes
_P = _c.c_void_p
_D = _c.c_double
_I = _c.c_int
_LL = _c.c_longlong

SIGNATURES = {
    "ib_plan_create": (_I, [_c.POINTER(ib_plan_desc), _c.POINTER(_P)]),
    "ib_plan_set_params": (_I, [_P, _I, _c.POINTER(_D), _c.POINTER(_D), _P]),
    "ib_plan_destroy": (_I, [_P]),
    "ib_plan_set_tuning": (_I, [_P, _LL, _LL, _LL]),
    "ib_plan_set_tiles": (_I, [_P, _I, _I]),
    "ib_plan_get_tiles": (_I, [_P, _c.POINTER(_I), _c.POINTER(_I)]),
    "ib_plan_set_timing": (_I, [_P, _I]),
    "ib_plan_get_timing": (_I, [_P, _c.POINTER(_D), _c.POINTER(_LL), _c.POINTER(_LL)]),
    "ib_collide": (_I, [_P, _P, _P, _LL, _P, _I, _D, _P]),
    "ib_collide_static": (_I, [_P, _P, _P, _LL, _P, _I, _D, _P]),
    "ib_collide_group": (_I, [_I, _c.POINTER(_P), _c.POINTER(_P), _c.POINTER(_P), _c.POINTER(_LL),
                              _c.POINTER(_P), _I, _D, _P]),
    "ib_set_group_mode": (_I, [_I]),
    "ib_laplacian": (_I, [_P, _P, _P, _LL, _P]),
    "ib_moments": (_I, [_P, _P, _LL, _I, _I, _D, _P]),
    "ib_project_conservative": (_I, [_P, _LL, _I, _I, _D, _P]),
    "ib_tvd1d": (_I, [_P, _P, _I, _LL, _P, _D, _I, _I, _I, _I, _I, _P]),
    "ib_tvd_axis": (_I, [_P, _P, _LL, _I, _LL, _P, _LL, _D, _I, _I, _I, _I, _I, _P]),
    "ib_upwind_flux": (_I, [_P, _P, _P, _LL, _LL, _D, _D, _I, _P]),
    "ib_fast_sweep": (_I, [_P, _P, _P, _P, _I, _I, _LL, _D, _D, _P, _P]),
    "ib_axpy": (_I, [_P, _P, _D, _LL, _P]),
    "ib_axpy_chain": (_I, [_P, _P, _D, _D, _D, _LL, _P]),
    "ib_tvd_limit": (_I, [_P, _P, _LL, _LL, _P, _LL, _I, _P]),
    "ib_plan_scratch_need": (_I, [_P, _LL, _I, _P, _P]),
    "ib_ghost_embed": (_I, [_P, _P, _I, _LL, _LL, _LL, _I, _I, _P]),
    "ib_ghost_fill": (_I, [_P, _LL, _LL, _LL, _I, _I, _I, _P]),
    "ib_df_combine": (_I, [_P, _P, _P, _D, _D, _LL, _P]),
    "ib_host_sp": (_D, [_I, _D]),
    "ib_last_error": (_c.c_char_p, []),
    "ib_version": (_c.c_char_p, []),
}

_lib = None


def load():
    """Load the HIP library (once).  Raises ImportError if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.environ.get("INELASTIC_HIP_LIB", str(LIB_PATH))
    if not os.path.exists(path):
        raise ImportError(
            "libinelastic_hip.so not found at %s; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)" % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    return load().ib_last_error().decode()


def check(rc, what=""):
    """Map a C status to the exception type the reference raises at that point."""
    if rc == IB_OK:
        return
    msg = last_error() or what
    if rc == IB_ERR_ARG:
        raise ValueError(msg)
    if rc == IB_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == IB_ERR_BC:
        raise Exception(msg)
    raise RuntimeError("HIP error in %s: %s" % (what, msg))


def dptr(a):
    """Array of doubles -> ctypes pointer (host, numpy)."""
    import numpy as np
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
