"""ctypes binding of libgammcor_hip.so.

The library is the product path: if it is missing or fails to load this
module raises, there is no fallback implementation.
"""
from __future__ import annotations

import atexit
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgammcor_hip.so")

_c_int, _c_i64, _c_double, _vp = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p

# name -> argtypes (all functions return int status unless noted)
SIGNATURES = {
    "gmc_version": [],
    "gmc_init": [],
    "gmc_finalize": [],
    "gmc_twono1": [_vp, _vp, _c_int, _c_i64],
    "gmc_twono1_dev": [_vp, _vp, _c_int, _c_i64, _vp],
    "gmc_ac0cas": [_vp, _vp, _vp, _c_i64, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int],
    "gmc_ac0cas_dev": [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _vp],
    "gmc_ac0cas_fofo_dev": [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _vp],
    "gmc_ac0cas_fofo": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int],
    "gmc_acecorr_dev": [_c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _vp, _c_int, _vp,
                        _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "gmc_acecorr_fofo_dev": [_c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _vp, _c_int,
                             _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "gmc_syevd_dev": [_c_int, _c_int, _c_int, _vp, _vp, _vp],
    "gmc_syevd_noalloc_dev": [_c_int, _c_int, _c_int, _vp, _vp, _vp],
    "gmc_acecorr_batch_dev": [_c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _vp, _vp, _vp,
                              _c_int, _vp, _vp, _vp, _vp],
    "gmc_batch_plan": [_c_int, _vp, _c_int, _vp, _vp, _vp],
    "gmc_sapt_response_dev": [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _vp,
                              _vp],
    "gmc_e2disp_dev": [_c_int, _c_int, _vp, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "gmc_read_aotwoint_molpro": [ctypes.c_char_p, _c_int, _vp],
    "gmc_read_rdm2_molpro": [ctypes.c_char_p, _c_int, _c_int, _c_int, _vp],
    "gmc_read_rdm1_molpro": [ctypes.c_char_p, _c_int, _c_int, _c_int, _vp],
    "gmc_read_mo_molpro": [ctypes.c_char_p, ctypes.c_char_p, _c_int, _vp],
    "gmc_read_oneint_molpro": [ctypes.c_char_p, ctypes.c_char_p, _c_int, _vp],
    "gmc_read_dpqrs": [ctypes.c_char_p, _c_int, _c_int, _vp],
    "gmc_prof_enable": [_c_int],
    "gmc_prof_reset": [],
    "gmc_prof_get": [ctypes.c_char_p, _vp, _vp, _vp],
    "gmc_ab_cas": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_i64, _vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp,
                   _c_double, _vp],
    "gmc_erpasymm1": [_vp, _vp, _vp, _vp, _c_int, _vp],
    "gmc_aceneerpa": [_vp, _vp, _vp, _vp, _c_i64, _vp, _vp, _vp, _vp, _c_int, _c_int],
    "gmc_aceinteg": [_c_int, _vp, _vp, _vp, _c_i64, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int,
                     _c_double, _vp, _vp, _vp, _vp],
    "gmc_aceinteg_fofo": [_c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int,
                          _c_double, _vp, _vp, _vp, _vp],
    "gmc_acecorr": [_c_int, _c_int, _c_int, _c_int, _c_int, _vp, _c_i64, _vp, _vp, _vp, _vp, _vp, _c_int, _vp, _c_int,
                    _vp, _vp, _vp, _vp, _vp, _vp],
    "gmc_forget_resident": [],
    "gmc_set_deterministic": [_c_int],
    "gmc_gemm_dev": [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _vp, _c_i64, _c_i64, _vp,
                     _c_i64, _c_i64, _c_double, _vp, _c_i64, _c_i64, _c_int, _vp],
    "gmc_gemm_lower_dev": [_c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _vp, _c_i64, _c_i64, _vp,
                           _c_i64, _c_i64, _c_double, _vp, _c_i64, _c_i64, _c_int, _vp],
    "gmc_potrf_dev": [_c_int, _c_int, _vp, _vp, _vp],
}

_lib = None


class GmcError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GmcError(f"{LIB_PATH} is not built; run `python -m gammcor_amd.build` "
                           "(there is no CPU fallback)")
        l = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(l, name)
            f.argtypes = args
            f.restype = _c_int
        l.gmc_last_error.argtypes = []
        l.gmc_last_error.restype = ctypes.c_char_p
        _lib = l
        # release the library's device memory, rocBLAS handle, events and
        # stream while the HIP runtime (and a profiler attached to it) is
        # still up: Python's exit handlers run before the runtime's teardown
        atexit.register(_finalize)
    return _lib


def _finalize():
    if _lib is not None:
        _lib.gmc_finalize()


def check(rc):
    if rc != 0:
        msg = lib().gmc_last_error().decode(errors="replace")
        raise GmcError(f"gammcor_hip status {rc}: {msg}")
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args))


def prof_get(name):
    """(total_ms, launches, total_work) of one instrumented kernel name."""
    t = ctypes.c_double()
    c = ctypes.c_int()
    w = ctypes.c_double()
    call("gmc_prof_get", name.encode(), ctypes.addressof(t), ctypes.addressof(c), ctypes.addressof(w))
    return t.value, c.value, w.value


def exported_symbols():
    return list(SIGNATURES) + ["gmc_last_error"]
