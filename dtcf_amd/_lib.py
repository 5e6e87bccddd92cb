"""ctypes binding of libdtcf.so (the C ABI in include/dtcf_local.h).

The library is built in-tree by ``__graft_entry__.build()`` (``make -C
dtcf_amd/csrc``: hipcc, gfx950).  If it is missing, importing the binding
raises: there is no CPU fallback anywhere in the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdtcf.so")
DROPIN_PATH = os.path.join(_HERE, "libdtcf_dropin.so")

DTCF_OK = 0
DTCF_MAX_DIMS = 8
STATUS = {
    0: "DTCF_OK",
    -1: "DTCF_ERR_INVALID",
    -2: "DTCF_ERR_UNSUPPORTED",
    -3: "DTCF_ERR_HIP",
    -4: "DTCF_ERR_NOMEM",
    -5: "DTCF_ERR_BUDGET",
    -6: "DTCF_ERR_NOTREADY",
}


class DtcfError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Strided(C.Structure):
    _fields_ = [
        ("dims", C.c_int32),
        ("ext", C.c_int64 * DTCF_MAX_DIMS),
        ("istr", C.c_int64 * DTCF_MAX_DIMS),
        ("ostr", C.c_int64 * DTCF_MAX_DIMS),
    ]


_i64 = C.c_int64
_vp = C.c_void_p
_lock = threading.Lock()
_L = None


def _declare(L):
    def fn(name, args, res=C.c_int):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
        return f

    fn("dtcf_last_error", [], C.c_char_p)
    fn("dtcf_init", [C.c_int])
    fn("dtcf_abi_version", [])
    fn("dtcf_permute_copy_f64", [_vp, _vp, C.POINTER(Strided), _i64, _vp, _i64, _vp, _i64,
                                 C.c_double, C.c_double, _vp])
    fn("dtcf_permute_tiles_f64", [_vp, _vp, C.c_int, C.POINTER(_i64), C.POINTER(C.c_int32), _i64,
                                  _vp])
    fn("dtcf_gemm_f64", [C.c_int, C.c_int, _i64, _i64, _i64, C.c_double, _vp, _i64, _vp, _i64,
                         _vp, _i64, _vp])
    fn("dtcf_view_create", [_i64, _i64, _vp, _vp, _vp, _i64, _i64, C.POINTER(_vp)])
    fn("dtcf_view_destroy", [_vp])
    fn("dtcf_view_span", [_vp], _i64)
    fn("dtcf_gemm_view_f64", [_i64, _i64, _i64, C.c_double, _vp, _vp, _vp, _vp, _vp, _vp, _i64,
                              _vp])
    fn("dtcf_view_gather_f64", [_vp, _vp, _vp, _i64, _vp])
    fn("dtcf_view_scatter_f64", [_vp, _i64, _vp, _vp, _vp])
    fn("dtcf_gemm_set_config", [C.c_int])
    fn("dtcf_gemm_f64_host", [C.c_int, C.c_int, _i64, _i64, _i64, C.c_double, _vp, _i64, _vp,
                              _i64, _vp, _i64])
    fn("dtcf_permute_tiles_f64_host", [_vp, _vp, C.c_int, C.POINTER(_i64), C.POINTER(C.c_int32),
                                       _i64])
    fn("dtcf_denom_update_f64", [_vp, _vp, _vp, _vp, _i64, _i64, _vp])
    fn("dtcf_energy_f64", [_vp, _vp, _i64, _i64, _vp, _vp])
    fn("dtcf_norm2_f64", [_vp, _i64, _vp, _vp])
    return L


def lib() -> C.CDLL:
    """Load libdtcf.so, raising if it has not been built."""
    global _L
    with _lock:
        if _L is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                    "g.build()'` (hipcc, gfx950); there is no CPU fallback")
            _L = _declare(C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL))
        return _L


def check(status: int) -> None:
    if status != DTCF_OK:
        raise DtcfError(status, lib().dtcf_last_error().decode(errors="replace"))


def i64_array(vals):
    arr = (_i64 * len(vals))(*[int(v) for v in vals])
    return arr


def i32_array(vals):
    return (C.c_int32 * len(vals))(*[int(v) for v in vals])
