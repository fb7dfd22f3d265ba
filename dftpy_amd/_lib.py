"""ctypes binding of ``libofdft_amd.so`` (C ABI declared in include/ofdft_amd.h).

There is no CPU fallback: importing the library fails with ImportError when
the shared object has not been built, and every call raises OfdftError when
the C side reports an error.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libofdft_amd.so")
HEADER_PATH = os.path.join(_HERE, "..", "include", "ofdft_amd.h")

F64 = 0
F32 = 1

E_TF, E_VW, E_NL, E_XC, E_H, E_PP, S_VRHO, S_RHO, S_DTHETA = range(9)
NSUM = 9

CONV_NONE, CONV_HARTREE, CONV_GG, CONV_WT = range(4)

# ofdft_tn_direction outcomes: the `stat` strings of get_direction_TN
TN_CONV, TN_WARN, TN_FAILED, TN_NOTREDUCE, TN_SMALL, TN_MAXFUN = range(6)
NSTRESS = 19  # OFDFT_NSTRESS
GGA_NOUT = 10  # OFDFT_GGA_NOUT

TN_STATUS = ("CONV", "WARN", "FAILED", "WARN : Not reduce", "WARN : Change too small", "NOTCONV")


class OfdftError(RuntimeError):
    """Error reported by the HIP library."""


class Functional(ctypes.Structure):
    _fields_ = [
        ("tf", ctypes.c_int), ("vw", ctypes.c_int), ("wt", ctypes.c_int),
        ("lda", ctypes.c_int), ("hartree", ctypes.c_int), ("pseudo", ctypes.c_int),
        ("x", ctypes.c_double), ("y", ctypes.c_double), ("alpha", ctypes.c_double),
        ("beta", ctypes.c_double), ("rho0", ctypes.c_double),
    ]


class VLine(ctypes.Structure):
    _fields_ = [
        ("knots", ctypes.POINTER(ctypes.c_double)), ("coefs", ctypes.POINTER(ctypes.c_double)),
        ("nknots", ctypes.c_int), ("q1", ctypes.c_double), ("dq", ctypes.c_double), ("qmax", ctypes.c_double),
        ("v0", ctypes.c_double), ("v1", ctypes.c_double), ("v2", ctypes.c_double),
    ]


class LwtTable(ctypes.Structure):
    """ofdft_lwt_table"""
    _fields_ = [
        ("kind", ctypes.c_int), ("knots", ctypes.POINTER(ctypes.c_double)),
        ("coefs", ctypes.POINTER(ctypes.c_double)), ("n", ctypes.c_int), ("x0", ctypes.c_double),
        ("dx", ctypes.c_double), ("etamax", ctypes.c_double), ("limit", ctypes.c_double),
    ]


class LwtParams(ctypes.Structure):
    """ofdft_lwt_params"""
    _fields_ = [
        ("alpha", ctypes.c_double), ("beta", ctypes.c_double), ("kfmax", ctypes.c_double),
        ("ld", ctypes.c_double), ("rhomax_ld", ctypes.c_double), ("kdd", ctypes.c_int),
        ("hermite", ctypes.c_int), ("mgpe", ctypes.c_int), ("mgpe_a", ctypes.c_double),
        ("mgpe_b", ctypes.c_double), ("mgpe_c", ctypes.c_double),
    ]


_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_ll = ctypes.c_longlong
_dp = ctypes.POINTER(ctypes.c_double)

_SIGS = {
    "ofdft_plan_create": (_i, [ctypes.POINTER(_vp), _i, _i, _i, _dp, _i, _i, _vp]),
    "ofdft_plan_destroy": (_i, [_vp]),
    "ofdft_plan_set_stream": (_i, [_vp, _vp]),
    "ofdft_plan_workspace_bytes": (_ll, [_vp]),
    "ofdft_set_functional": (_i, [_vp, ctypes.POINTER(Functional)]),
    "ofdft_set_vloc": (_i, [_vp, _vp]),
    "ofdft_eval": (_i, [_vp, _vp, _vp, _dp]),
    "ofdft_eval_phi": (_i, [_vp, _vp, _vp, _d, _d, _vp, _vp, _dp]),
    "ofdft_hvp": (_i, [_vp, _vp, _vp, _vp, _vp, _d, _vp, _dp]),
    "ofdft_hvp_prepare": (_i, [_vp, _vp, _vp, _vp, _d, _vp, _vp, _dp]),
    "ofdft_hvp_fast": (_i, [_vp, _vp, _vp, _vp, _vp, _dp]),
    "ofdft_potential_jvp": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "ofdft_tn_inner": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _dp]),
    "ofdft_tn_direction": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), _dp]),
    "ofdft_fft_r2c": (_i, [_vp, _vp, _vp]),
    "ofdft_fft_c2r": (_i, [_vp, _vp, _vp]),
    "ofdft_convolve": (_i, [_vp, _i, _vp, _vp]),
    "ofdft_dot3": (_i, [_vp, _vp, _vp, _dp]),
    "ofdft_axpby": (_i, [_vp, _d, _vp, _d, _vp]),
    "ofdft_lincomb": (_i, [_vp, _d, _vp, _d, _vp, _vp]),
    "ofdft_residual": (_i, [_vp, _vp, _vp, _d, _vp, _dp]),
    "ofdft_cg_update": (_i, [_vp, _d, _vp, _vp, _vp, _vp, _dp]),
    "ofdft_dot_vphip": (_i, [_vp, _vp, _vp, _vp, _dp]),
    "ofdft_orth_sums": (_i, [_vp, _vp, _vp, _vp, _dp]),
    "ofdft_orth_normalize": (_i, [_vp, _vp, _vp, _vp]),
    "ofdft_hvp_prepare_async": (_i, [_vp, _vp, _vp, _vp, _d, _vp, _vp]),
    "ofdft_plan_deferred_sums": (_i, [_vp, _dp]),
    "ofdft_convert": (_i, [_vp, _vp, _i, _vp, _ll]),
    "ofdft_vloc_build": (_i, [_vp, _i, ctypes.POINTER(VLine), _i, _dp, ctypes.POINTER(ctypes.c_int), _i, _i, _vp]),
    "ofdft_local_forces": (_i, [_vp, _vp, _dp]),
    "ofdft_ewald": (_i, [_vp, _i, _dp, _dp, _d, _d, _i, _i, _dp, _dp]),
    "ofdft_ewald_forces": (_i, [_vp, _i, _dp, _dp, _d, _d, _i, _i, _dp]),
    "ofdft_ewald_noalloc": (_i, [_vp, _i, _dp, _dp, _d, _d, _i, _i, _dp, _dp]),
    "ofdft_functional_stress": (_i, [_vp, _vp, _d, _dp]),
    "ofdft_local_stress": (_i, [_vp, _vp, _dp]),
    "ofdft_ewald_stress": (_i, [_vp, _i, _dp, _dp, _d, _d, _i, _i, _dp, _dp]),
    "ofdft_gga": (_i, [_vp, _i, _dp, _i, _vp, _vp, _vp, _d, _d, _vp, _dp]),
    "ofdft_lwt_set_kernel": (_i, [_vp, ctypes.POINTER(LwtTable), ctypes.POINTER(LwtTable)]),
    "ofdft_lwt_range": (_i, [_vp, _vp, _vp, _vp, _d, _d, _dp]),
    "ofdft_lwt": (_i, [_vp, ctypes.POINTER(LwtParams), _dp, _i, _vp, _vp, _vp, _d, _d, _vp, _dp]),
    "ofdft_plan_set_profiling": (_i, [_vp, _i]),
    "ofdft_plan_set_graph": (_i, [_vp, _i]),
    "ofdft_plan_set_tiles": (_i, [_vp, _i, _i]),
    "ofdft_plan_set_cg_spectra": (_i, [_vp, _i, _i, _i]),
    "ofdft_plan_cg_spectra": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "ofdft_batch_create": (_i, [ctypes.POINTER(_vp), ctypes.POINTER(_vp), _i]),
    "ofdft_batch_destroy": (_i, [_vp]),
    "ofdft_batch_set_graph": (_i, [_vp, _i]),
    "ofdft_batch_tn_direction": (_i, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                      ctypes.POINTER(_vp), ctypes.POINTER(_vp), _i, ctypes.POINTER(_i),
                                      ctypes.POINTER(_i), _dp]),
    "ofdft_plan_get_profile": (_i, [_vp, _dp, ctypes.POINTER(_ll), _dp]),
    "ofdft_last_error": (ctypes.c_char_p, []),
    "ofdft_version": (_i, []),
}

_lib = None


def load():
    """Load the HIP library once; ImportError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libofdft_amd.so is not built (expected at %s); run `python -m dftpy_amd.build_lib`" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = load().ofdft_last_error().decode(errors="replace")
        raise OfdftError("ofdft error %d: %s" % (rc, msg))
    return rc


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def darr(n):
    return (ctypes.c_double * n)()


def exported_symbols():
    """Function names declared in include/ofdft_amd.h."""
    import re
    with open(HEADER_PATH) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"\b(ofdft_[a-z0-9_]+)\s*\(", text)))


def as_numpy(carr):
    return np.array(carr[:], dtype=np.float64)
