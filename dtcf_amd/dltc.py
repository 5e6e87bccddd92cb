"""DLTC tile contractions on the device (wraps the dtcf_dltc_* C ABI).

Index names follow DLTC/param.h (non-CHECK build): h1..h12 -> 0..11,
p1..p12 -> 12..23; names below ``nocc`` are occupied.  A ``Term`` is
``tC += coef * tA * tB`` (DLTC/expr.cc:13-67); ``Term.task`` runs one task
over a tile box as DLTC/task.cc:83-137 does, ``Term.run`` the whole term.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import check, lib

NIDX = 24
NAMES = [f"h{i}" for i in range(1, 13)] + [f"p{i}" for i in range(1, 13)]
IDX = {n: i for i, n in enumerate(NAMES)}


class Space(C.Structure):
    _fields_ = [("nocc", C.c_int32), ("O", C.c_int32), ("V", C.c_int32), ("wbO", C.c_int32),
                ("wbV", C.c_int32)]


class TensorDesc(C.Structure):
    _fields_ = [("dims", C.c_int32), ("ids", C.c_int32 * _lib.DTCF_MAX_DIMS), ("data", C.c_void_p)]


def ids_of(names):
    return [IDX[n] if isinstance(n, str) else int(n) for n in names]


def make_space(O, V, wbO, wbV, nocc=12):
    return Space(nocc, O, V, wbO, wbV)


def desc(ids, data: torch.Tensor):
    d = TensorDesc()
    d.dims = len(ids)
    for i, x in enumerate(ids):
        d.ids[i] = x
    d.data = data.data_ptr()
    return d


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _declare():
    L = lib()
    vp = C.c_void_p
    sig = {
        "dtcf_dltc_mul_create": [vp, vp, vp, vp, C.c_double, C.POINTER(vp)],
        "dtcf_dltc_mul_info": [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                               C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
        "dtcf_dltc_mul_task": [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), vp],
        "dtcf_dltc_mul_run": [vp, vp],
        "dtcf_dltc_mul_destroy": [vp],
        "dtcf_dltc_sum": [vp, vp, vp, C.c_double, C.POINTER(C.c_int32), C.POINTER(C.c_int32), vp],
        "dtcf_dltc_transpose_f64": [vp, C.POINTER(C.c_int32), vp, C.POINTER(C.c_int32), C.c_int,
                                    vp, vp],
        "dtcf_view_sub": [vp, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.POINTER(vp)],
        "dtcf_gemm_reserve_workspace": [vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int


_declare()


class Term:
    """tC += coef * tA * tB over dense device tensors."""

    def __init__(self, space: Space, c_ids, C_t, a_ids, A_t, b_ids, B_t, coef=1.0):
        self.space = space
        self._keep = (C_t, A_t, B_t)
        self.dc, self.da, self.db = (desc(ids_of(c_ids), C_t), desc(ids_of(a_ids), A_t),
                                     desc(ids_of(b_ids), B_t))
        h = C.c_void_p()
        check(lib().dtcf_dltc_mul_create(C.byref(space), C.byref(self.dc), C.byref(self.da),
                                         C.byref(self.db), float(coef), C.byref(h)))
        self.h = h
        mnk = (C.c_int64 * 3)()
        mi, ni, ki = (C.c_int32 * NIDX)(), (C.c_int32 * NIDX)(), (C.c_int32 * NIDX)()
        cnt = (C.c_int32 * 3)()
        check(lib().dtcf_dltc_mul_info(h, mnk, mi, ni, ki, cnt))
        self.mnk = tuple(mnk)
        self.m_ids = list(mi[:cnt[0]])
        self.n_ids = list(ni[:cnt[1]])
        self.k_ids = list(ki[:cnt[2]])

    @property
    def flops(self):
        M, N, K = self.mnk
        return 2.0 * M * N * K

    def run(self):
        check(lib().dtcf_dltc_mul_run(self.h, _stream()))

    def task(self, itlb, itub):
        lb = (C.c_int32 * NIDX)(*itlb)
        ub = (C.c_int32 * NIDX)(*itub)
        check(lib().dtcf_dltc_mul_task(self.h, lb, ub, _stream()))

    def __del__(self):
        try:
            if self.h:
                lib().dtcf_dltc_mul_destroy(self.h)
                self.h = None
        except Exception:
            pass


class TaskGraph:
    """Tile tasks recorded once as a HIP graph and replayed as one launch.

    DLTC issues one task per tile box (DLTC/task.cc:83-137, driven by the
    loops of DLTC/ccsd-t2.cc); a task is a handful of small GEMMs, each a
    kernel of a few microseconds, and issued one by one from the host they
    cost about 10 us each.  `chains` is a list of callables, each enqueueing
    on the current stream tasks that must run in order (the sum tiles of one
    output tile); different chains must write disjoint outputs.  Under
    capture the chains are spread over `branches` streams forked from and
    joined back into the capturing stream, so the graph has that many
    independent branches; every replay() re-runs the recorded kernels with
    the same buffers, which must stay allocated and in place.  Per output
    tile the order of the tasks is the eager one, so results are bitwise
    those of issuing the chains one after the other.  Call each entry point
    once eagerly before capturing (first calls set kernel attributes).  A
    GEMM that splits K uses its stream's workspace, which cannot grow under
    capture: every branch stream's workspace is reserved first (at the bound
    any GEMM can need), so the graph never allocates and later growth on a
    reused stream cannot free a buffer the graph holds."""

    def __init__(self, chains, branches=4):
        import torch
        self.graph = torch.cuda.CUDAGraph()
        self.streams = [torch.cuda.Stream() for _ in range(max(1, branches))]
        for st in self.streams:
            check(lib().dtcf_gemm_reserve_workspace(C.c_void_p(st.cuda_stream)))
        with torch.cuda.graph(self.graph):
            cur = torch.cuda.current_stream()
            fork = torch.cuda.Event()
            fork.record(cur)
            for st in self.streams:
                st.wait_event(fork)
            for i, chain in enumerate(chains):
                with torch.cuda.stream(self.streams[i % len(self.streams)]):
                    chain()
            for st in self.streams:
                join = torch.cuda.Event()
                join.record(st)
                cur.wait_event(join)

    def replay(self):
        self.graph.replay()


def task_sum(space, c_ids, C_t, a_ids, A_t, coef=1.0, itlb=None, itub=None):
    dc, da = desc(ids_of(c_ids), C_t), desc(ids_of(a_ids), A_t)
    lb = None if itlb is None else (C.c_int32 * NIDX)(*itlb)
    ub = None if itub is None else (C.c_int32 * NIDX)(*itub)
    check(lib().dtcf_dltc_sum(C.byref(space), C.byref(dc), C.byref(da), float(coef), lb, ub,
                              _stream()))


def transpose(buf_x, x_ids, buf_y, y_ids, space):
    x = (C.c_int32 * len(x_ids))(*ids_of(x_ids))
    y = (C.c_int32 * len(y_ids))(*ids_of(y_ids))
    check(lib().dtcf_dltc_transpose_f64(C.c_void_p(buf_x.data_ptr()), x,
                                        C.c_void_p(buf_y.data_ptr()), y, len(x_ids),
                                        C.byref(space), _stream()))


def tile_counts(space, ids):
    """Tiles per index name (getSP, DLTC/define.h:67-73)."""
    return [space.O // space.wbO if i < space.nocc else space.V // space.wbV for i in ids]
