"""Torch-level wrappers of the C ABI (device tensors in, device tensors out).

Every call goes through libdtcf.so; tensors must be float64 CUDA (HIP)
tensors on the current device.  Work is enqueued on torch's current stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, lib


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: torch.Tensor | None):
    if t is None:
        return None
    assert t.is_cuda, "device tensor expected"
    return C.c_void_p(t.data_ptr())


def _f64(t: torch.Tensor):
    assert t.dtype == torch.float64 and t.is_cuda and t.is_contiguous(), \
        "contiguous float64 device tensor expected"
    return _ptr(t)


# ---------------------------------------------------------------- permutes
def permute_copy(inp, out, ext, istr, ostr, ntiles=1, src_off=None, src_step=0, dst_off=None,
                 dst_step=0, alpha=1.0, beta=0.0):
    d = _lib.Strided()
    d.dims = len(ext)
    for i, (e, a, b) in enumerate(zip(ext, istr, ostr)):
        d.ext[i], d.istr[i], d.ostr[i] = int(e), int(a), int(b)
    check(lib().dtcf_permute_copy_f64(_f64(inp), _f64(out), C.byref(d), int(ntiles),
                                      _ptr(src_off), int(src_step), _ptr(dst_off), int(dst_step),
                                      float(alpha), float(beta), _stream()))
    return out


def permute_tiles(inp, widths, out_src, ntiles=1, out=None):
    """ntiles dense row-major tiles of extents `widths`; output dim n is input
    dim out_src[n] (DLTC convention)."""
    if out is None:
        out = torch.empty_like(inp)
    w = _lib.i64_array(widths)
    s = _lib.i32_array(out_src)
    check(lib().dtcf_permute_tiles_f64(_f64(inp), _f64(out), len(widths), w, s, int(ntiles),
                                       _stream()))
    return out


def rrr_transpose(inp, pmap, widths, ntiles=1, out=None):
    """RRR convention (helper.cc:275-522): input dim i -> output position pmap[i]."""
    out_src = [0] * len(pmap)
    for i, p in enumerate(pmap):
        out_src[p] = i
    return permute_tiles(inp, widths, out_src, ntiles, out)


# ---------------------------------------------------------------- GEMM
def gemm(transa, transb, m, n, k, alpha, A, lda, B, ldb, Cm, ldc):
    check(lib().dtcf_gemm_f64(int(transa), int(transb), int(m), int(n), int(k), float(alpha),
                              _f64(A), int(lda), _f64(B), int(ldb), _f64(Cm), int(ldc),
                              _stream()))
    return Cm


def set_gemm_config(cfg: int):
    check(lib().dtcf_gemm_set_config(int(cfg)))


class View:
    """Owning handle of a dtcf_view (offset tables on the device)."""

    def __init__(self, rows, cols, rofs, cofs, blk=None, rpb=0, cpb=0):
        r = np.ascontiguousarray(rofs, dtype=np.int64)
        c = np.ascontiguousarray(cofs, dtype=np.int64)
        b = None if blk is None else np.ascontiguousarray(blk, dtype=np.int64)
        h = C.c_void_p()
        check(lib().dtcf_view_create(int(rows), int(cols), r.ctypes.data_as(C.c_void_p),
                                     c.ctypes.data_as(C.c_void_p),
                                     None if b is None else b.ctypes.data_as(C.c_void_p),
                                     int(rpb), int(cpb), C.byref(h)))
        self.h = h
        self.rows, self.cols = int(rows), int(cols)

    @property
    def span(self):
        return int(lib().dtcf_view_span(self.h))

    def __del__(self):
        try:
            if self.h:
                lib().dtcf_view_destroy(self.h)
                self.h = None
        except Exception:
            pass


def gemm_view(m, n, k, alpha, A, av: View, B, bv: View, Cm, cv: View | None = None, ldc=0):
    check(lib().dtcf_gemm_view_f64(int(m), int(n), int(k), float(alpha), _f64(A), av.h, _f64(B),
                                   bv.h, _f64(Cm), None if cv is None else cv.h, int(ldc),
                                   _stream()))
    return Cm


def view_gather(src, v: View, dst, ld):
    check(lib().dtcf_view_gather_f64(_f64(src), v.h, _f64(dst), int(ld), _stream()))
    return dst


def view_scatter(src, ld, v: View, dst):
    check(lib().dtcf_view_scatter_f64(_f64(src), int(ld), v.h, _f64(dst), _stream()))
    return dst


# ---------------------------------------------------------------- reductions
def denom_update(R, eo, ev, T=None):
    o, v = eo.numel(), ev.numel()
    if T is None:
        T = torch.empty_like(R)
    check(lib().dtcf_denom_update_f64(_f64(R), _f64(T), _f64(eo), _f64(ev), o, v, _stream()))
    return T


def energy(V, T, o, v, out=None):
    if out is None:
        out = torch.zeros(1, dtype=torch.float64, device=V.device)
    check(lib().dtcf_energy_f64(_f64(V), _f64(T), int(o), int(v), _f64(out), _stream()))
    return out


def norm2(R, out=None):
    if out is None:
        out = torch.zeros(1, dtype=torch.float64, device=R.device)
    check(lib().dtcf_norm2_f64(_f64(R), R.numel(), _f64(out), _stream()))
    return out
