"""Mixed-precision contractions with FP64 recombination (wraps
dtcf_split_f64 / dtcf_gemm_split_f64) and device-side synthetic inputs."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import ops
from ._lib import check, lib

F32, BF16X3, F16X3, I8 = 1, 2, 3, 4
NAMES = {F32: "fp32", BF16X3: "bf16x3", F16X3: "fp16x3", I8: "int8"}


def _declare():
    L = lib()
    vp, i64 = C.c_void_p, C.c_int64
    L.dtcf_split_bytes.argtypes = [C.c_int, C.c_int, i64, i64]
    L.dtcf_split_bytes.restype = i64
    L.dtcf_split_f64.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp]
    L.dtcf_gemm_split_f64.argtypes = [C.c_int, C.c_int, i64, i64, i64, C.c_double, vp, vp, vp, vp,
                                      C.c_int, vp, vp, i64, vp]
    L.dtcf_fill_normal_f64.argtypes = [vp, i64, C.c_uint64, C.c_double, i64, vp]
    L.dtcf_fill_uniform_f64.argtypes = [vp, i64, C.c_uint64, C.c_double, C.c_double, vp]
    L.dtcf_antisymmetrize4_f64.argtypes = [vp, vp, i64, i64, i64, i64, vp]
    L.dtcf_gemm_split_set_fused.argtypes = [C.c_int]
    for f in ("dtcf_split_f64", "dtcf_gemm_split_f64", "dtcf_fill_normal_f64", "dtcf_gemm_split_set_fused",
              "dtcf_fill_uniform_f64", "dtcf_antisymmetrize4_f64"):
        getattr(L, f).restype = C.c_int


_declare()


def _s():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def set_fused(on: bool):
    """Slice products (INT8 digits; three BF16 / FP16 slices at the default
    level 4) in one fused kernel (default) or one GEMM each."""
    check(lib().dtcf_gemm_split_set_fused(1 if on else 0))


def fill_normal(n, seed, sigma=1.0, device="cuda", start=0):
    out = torch.empty(n, dtype=torch.float64, device=device)
    check(lib().dtcf_fill_normal_f64(_p(out), n, seed, sigma, start, _s()))
    return out


def fill_uniform(n, seed, lo, hi, device="cuda"):
    out = torch.empty(n, dtype=torch.float64, device=device)
    check(lib().dtcf_fill_uniform_f64(_p(out), n, seed, lo, hi, _s()))
    return out


def antisymmetrize4(x, shape):
    out = torch.empty_like(x)
    check(lib().dtcf_antisymmetrize4_f64(_p(x), _p(out), *[int(s) for s in shape], _s()))
    return out


def dense_view(rows, cols, ld=None, trans=False):
    """View of a dense row-major matrix (trans: element (r, c) at c*ld + r)."""
    ld = cols if ld is None and not trans else (rows if ld is None else ld)
    if trans:
        return ops.View(rows, cols, np.arange(rows), np.arange(cols) * ld)
    return ops.View(rows, cols, np.arange(rows) * ld, np.arange(cols))


class Split:
    """Slices of one FP64 operand (a view, rows x K)."""

    def __init__(self, prec, src, view: ops.View, nslice=4):
        self.prec, self.nslice = prec, (nslice if prec == I8 else (1 if prec == F32 else 3))
        self.rows, self.K = view.rows, view.cols
        nbytes = lib().dtcf_split_bytes(prec, self.nslice, self.rows, self.K)
        self.buf = torch.empty(nbytes, dtype=torch.uint8, device=src.device)
        rp = (self.rows + 127) // 128 * 128
        self.rexp = (torch.zeros(rp, dtype=torch.int32, device=src.device)
                     if prec in (F16X3, I8) else None)
        check(lib().dtcf_split_f64(prec, self.nslice, _p(src), view.h, _p(self.buf),
                                   _p(self.rexp), _s()))


def gemm_split(a: Split, b: Split, Cm, alpha=1.0, cview: ops.View | None = None, ldc=0,
               max_level=0):
    assert a.prec == b.prec and a.K == b.K and a.nslice == b.nslice
    check(lib().dtcf_gemm_split_f64(a.prec, a.nslice, a.rows, b.rows, a.K, float(alpha),
                                    _p(a.buf), _p(a.rexp), _p(b.buf), _p(b.rexp), int(max_level),
                                    _p(Cm), None if cview is None else cview.h, int(ldc), _s()))
    return Cm


def n_products(prec, nslice=4, max_level=0):
    ns = 1 if prec == F32 else (nslice if prec == I8 else 3)
    lev = max_level if max_level > 0 else ns + 1
    return sum(1 for i in range(1, ns + 1) for j in range(1, ns + 1) if i + j <= lev)
