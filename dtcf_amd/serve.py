"""Serving CC iteration programs from device-resident tensors (wraps the
``dtcf_server_*`` / ``dtcf_program_*`` / ``dtcf_request_*`` C ABI).

A ``System`` holds one synthetic UHF system's input tensors (integrals, Fock
blocks, amplitudes) in the server's arena for the server's life; a
``Program`` is one term list (T1 or T2) planned over that system; each
``submit`` is a request that evaluates the program into fresh output
buffers.  The server admits requests by HBM budget and interleaves their
statements over its streams.

Systems are spin-orbital counts o, v split evenly into the alpha and beta
ranges of the term lists: Oa = Ob = o/2, Va = Vb = v/2.
"""
from __future__ import annotations

import ctypes as C
import zlib

import numpy as np
import torch

from . import mixed, termlist
from ._lib import check, lib

ROLE_INPUT, ROLE_TEMP, ROLE_OUTPUT = 0, 1, 2


class RequestInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("ntasks", C.c_int64), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double), ("flops", C.c_double)]


def _declare():
    L = lib()
    vp, i64, dbl = C.c_void_p, C.c_int64, C.c_double
    sig = {
        "dtcf_server_open": [i64, C.c_int, C.POINTER(vp)],
        "dtcf_server_close": [vp],
        "dtcf_server_alloc": [vp, i64, C.POINTER(vp)],
        "dtcf_server_free": [vp, vp],
        "dtcf_server_usage": [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)],
        "dtcf_server_inject_failure": [vp, i64],
        "dtcf_program_create": [vp, C.c_int, C.c_int, C.c_int, C.POINTER(vp)],
        "dtcf_program_array": [vp, C.c_char_p, C.c_int, C.POINTER(i64), C.c_int],
        "dtcf_program_bind": [vp, C.c_char_p, vp],
        "dtcf_program_contract": [vp, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                  C.c_char_p, dbl],
        "dtcf_program_sum": [vp, C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_char_p),
                             C.POINTER(C.c_char_p), C.POINTER(dbl)],
        "dtcf_program_finalize": [vp],
        "dtcf_program_set_layout": [vp, C.c_int],
        "dtcf_program_set_mixed_min": [vp, C.c_double],
        "dtcf_program_set_resident": [vp, C.c_char_p, C.c_int],
        "dtcf_program_get_layout": [vp, C.POINTER(C.c_int)],
        "dtcf_program_info": [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(dbl), C.POINTER(i64)],
        "dtcf_program_statement_plan": [vp, i64, C.POINTER(i64)],
        "dtcf_program_destroy": [vp],
        "dtcf_request_submit": [vp, vp, C.POINTER(vp), C.POINTER(i64)],
        "dtcf_request_wait": [vp, i64, dbl],
        "dtcf_request_info_get": [vp, i64, C.POINTER(RequestInfo)],
        "dtcf_request_task_times": [vp, i64, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl), i64],
        "dtcf_request_release": [vp, i64],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int


_declare()

_HIP = None


def _hip():
    """The HIP runtime, for raw device<->host copies of arena memory."""
    global _HIP
    if _HIP is None:
        _HIP = C.CDLL("libamdhip64.so")
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _HIP.hipMemcpy.restype = C.c_int
    return _HIP


def to_host(ptr, n):
    out = np.empty(n, dtype=np.float64)
    if n:
        st = _hip().hipMemcpy(out.ctypes.data, C.c_void_p(ptr), n * 8, 2)
        if st != 0:
            raise RuntimeError(f"hipMemcpy device->host failed ({st})")
    return out


def from_host(ptr, arr):
    arr = np.ascontiguousarray(arr, dtype=np.float64)
    if arr.size:
        st = _hip().hipMemcpy(C.c_void_p(ptr), arr.ctypes.data, arr.size * 8, 1)
        if st != 0:
            raise RuntimeError(f"hipMemcpy host->device failed ({st})")


def copy_device(dst, src, n):
    """n doubles device -> device (raw addresses)."""
    if n:
        st = _hip().hipMemcpy(C.c_void_p(dst), C.c_void_p(src), n * 8, 3)
        if st != 0:
            raise RuntimeError(f"hipMemcpy device->device failed ({st})")


def system_sizes(o, v):
    assert o % 2 == 0 and v % 2 == 0, "spin-orbital counts split evenly into alpha and beta"
    return {"Oa": o // 2, "Ob": o // 2, "Va": v // 2, "Vb": v // 2}


class Server:
    def __init__(self, budget_bytes: int, nslots: int = 4):
        torch.cuda.init()
        h = C.c_void_p()
        check(lib().dtcf_server_open(int(budget_bytes), int(nslots), C.byref(h)))
        self.h = h
        self.nslots = nslots

    def alloc(self, nelems):
        p = C.c_void_p()
        check(lib().dtcf_server_alloc(self.h, int(nelems) * 8, C.byref(p)))
        return p.value

    def free(self, ptr):
        check(lib().dtcf_server_free(self.h, C.c_void_p(ptr)))

    def usage(self):
        v = [C.c_int64() for _ in range(4)]
        check(lib().dtcf_server_usage(self.h, *[C.byref(x) for x in v]))
        return dict(budget=v[0].value, resident=v[1].value, reserved=v[2].value,
                    peak_reserved=v[3].value)

    def inject_failure(self, after):
        """The statement issued after `after` more fails on the host (tests)."""
        check(lib().dtcf_server_inject_failure(self.h, int(after)))

    def wait(self, rid, timeout_ms=-1.0):
        return lib().dtcf_request_wait(self.h, int(rid), float(timeout_ms))

    def info(self, rid):
        inf = RequestInfo()
        check(lib().dtcf_request_info_get(self.h, int(rid), C.byref(inf)))
        return inf

    def task_times(self, rid, n):
        a, b, c = (np.empty(n) for _ in range(3))
        dp = C.POINTER(C.c_double)
        check(lib().dtcf_request_task_times(self.h, int(rid), a.ctypes.data_as(dp),
                                            b.ctypes.data_as(dp), c.ctypes.data_as(dp), n))
        return a, b, c

    def release(self, rid):
        check(lib().dtcf_request_release(self.h, int(rid)))

    def close(self):
        if self.h:
            check(lib().dtcf_server_close(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class System:
    """Input tensors of one synthetic system, resident in the server arena.

    Each input is N(0, 0.05^2) from the device generator with a seed derived
    from the system seed and the array name; arrays declared with two
    antisymmetric pairs (0,1)(2,3) over equal extents are antisymmetrised."""

    def __init__(self, server: Server, o: int, v: int, lists, seed=7, sizes=None):
        self.server, self.o, self.v = server, o, v
        self.sizes = sizes or system_sizes(o, v)
        self.seed = seed
        self.ptr, self.shape = {}, {}
        names = []
        for tl in lists:
            for n in tl.inputs():
                if n not in names:
                    names.append(n)
        decl = {}
        for tl in lists:
            decl.update({n: tl.arrays[n] for n in names if n in tl.arrays})
        self.decl = decl
        for n in names:
            a = decl[n]
            shp = tuple(self.sizes[r] for r in a.ranges)
            cnt = int(np.prod(shp))
            p = server.alloc(cnt)
            raw = self.make(n, a, shp)
            torch.cuda.synchronize()
            st = _hip().hipMemcpy(C.c_void_p(p), C.c_void_p(raw.data_ptr()), cnt * 8, 3)
            if st != 0:
                raise RuntimeError(f"hipMemcpy device->device failed ({st})")
            del raw
            self.ptr[n], self.shape[n] = p, shp
        torch.cuda.synchronize()

    def name_seed(self, name):
        return (self.seed * 1000003 + zlib.crc32(name.encode())) & 0xFFFFFFFF

    def make(self, name, decl, shp):
        """Contents of input ``name`` (a flat float64 device tensor)."""
        cnt = int(np.prod(shp))
        raw = mixed.fill_normal(cnt, self.name_seed(name), 0.05)
        anti = len(shp) == 4 and set(decl.groups) == {(0, 1), (2, 3)} and \
            shp[0] == shp[1] and shp[2] == shp[3]
        if anti:
            raw = mixed.antisymmetrize4(raw, shp)
        return raw

    def host(self, name):
        return to_host(self.ptr[name], int(np.prod(self.shape[name]))).reshape(self.shape[name])

    def bytes(self):
        return sum(int(np.prod(s)) * 8 for s in self.shape.values())


LAYOUT_AUTO, LAYOUT_REUSE, LAYOUT_SEPARATE = 0, 1, 2


def integral_inputs(tl: termlist.TermList):
    """Inputs of a term list that stay the same in every CC iteration: the
    integrals and Fock blocks (every input but the amplitudes t*)."""
    return [n for n in tl.inputs() if not n.startswith("t")]


class Program:
    """One term list planned over a system.  ``expose`` makes every assigned
    array an output (for parity tests of each statement); ``layout`` picks the
    placement of temporaries (LAYOUT_*; ``self.layout`` is the one used);
    ``resident`` names inputs whose mixed-precision slices are made once at
    finalize (inputs that do not change between requests, e.g.
    ``integral_inputs(tl)``)."""

    def __init__(self, server: Server, tl: termlist.TermList, system: System, prec=0, nslice=5,
                 max_level=0, expose=False, layout=LAYOUT_AUTO, mixed_min=None, resident=()):
        self.server, self.tl, self.system = server, tl, system
        L = lib()
        h = C.c_void_p()
        check(L.dtcf_program_create(server.h, int(prec), int(nslice), int(max_level), C.byref(h)))
        self.h = h
        if mixed_min is not None:
            check(L.dtcf_program_set_mixed_min(h, float(mixed_min)))
        self.prec = prec
        used = []
        for s in tl.statements:
            for r in [s.lhs] + s.terms:
                if r.name not in used:
                    used.append(r.name)
        outs = set(tl.outputs())
        inputs = set(tl.inputs())
        self.outputs, self.out_shapes = [], []
        try:
            for n in used:
                shp = tuple(system.sizes[r] for r in tl.arrays[n].ranges)
                role = ROLE_INPUT if n in inputs else (
                    ROLE_OUTPUT if (n in outs or expose) else ROLE_TEMP)
                dims = (C.c_int64 * len(shp))(*shp)
                check(L.dtcf_program_array(h, n.encode(), len(shp), dims, role))
                if role == ROLE_INPUT:
                    check(L.dtcf_program_bind(h, n.encode(), C.c_void_p(system.ptr[n])))
                    if n in resident:
                        check(L.dtcf_program_set_resident(h, n.encode(), 1))
                elif role == ROLE_OUTPUT:
                    self.outputs.append(n)
                    self.out_shapes.append(shp)
            for s in tl.statements:
                ci = ",".join(s.lhs.idx).encode()
                if s.kind == "mul":
                    a, b = s.terms
                    check(L.dtcf_program_contract(h, s.lhs.name.encode(), ci, a.name.encode(),
                                                  ",".join(a.idx).encode(), b.name.encode(),
                                                  ",".join(b.idx).encode(), 1.0))
                else:
                    nt = len(s.terms)
                    names = (C.c_char_p * nt)(*[t.name.encode() for t in s.terms])
                    idx = (C.c_char_p * nt)(*[",".join(t.idx).encode() for t in s.terms])
                    coefs = (C.c_double * nt)(*[t.coef for t in s.terms])
                    check(L.dtcf_program_sum(h, s.lhs.name.encode(), ci, nt, names, idx, coefs))
            check(L.dtcf_program_set_layout(h, int(layout)))
            check(L.dtcf_program_finalize(h))
        except Exception:
            L.dtcf_program_destroy(h)
            self.h = None
            raise
        blk, nst, nout = C.c_int64(), C.c_int64(), C.c_int64()
        fl = C.c_double()
        check(L.dtcf_program_info(h, C.byref(blk), C.byref(nst), C.byref(fl), C.byref(nout)))
        self.block_bytes, self.nstatements, self.flops = blk.value, nst.value, fl.value
        assert nout.value == len(self.outputs)
        lay = C.c_int()
        check(L.dtcf_program_get_layout(h, C.byref(lay)))
        self.layout = lay.value

    PLAN_FLAGS = {1: "swap", 2: "mixed", 4: "pack_a", 8: "pack_b", 16: "a_colpairs",
                  32: "a_rowpairs", 64: "b_colpairs", 128: "b_rowpairs", 256: "c_dense",
                  512: "res_a", 1024: "res_b", 2048: "sum_plane"}
    TILES = {-1: "mixed", 0: "128x128", 1: "192x96", 2: "96x192", 3: "64x64", 4: "192x64", 5: "64x192",
             6: "thin"}

    def plan(self, i):
        """The finalized plan of statement i (dtcf_program_statement_plan):
        kind, GEMM m, n, k, tile configuration and access flags, or the
        number of terms of a sum."""
        out = (C.c_int64 * 8)()
        check(lib().dtcf_program_statement_plan(self.h, int(i), out))
        flags = [n for b, n in self.PLAN_FLAGS.items() if out[4] & b]
        if out[0] == 1:
            return {"kind": "sum", "nterms": out[6], "flags": flags}
        return {"kind": "mul", "m": out[1], "n": out[2], "k": out[3], "flags": flags,
                "tile": self.TILES.get(out[5], out[5])}

    def new_outputs(self, device="cuda"):
        return [torch.empty(int(np.prod(s)), dtype=torch.float64, device=device)
                for s in self.out_shapes]

    def submit(self, outputs=None):
        """Queues one evaluation; returns (request id, output tensors)."""
        outs = outputs if outputs is not None else self.new_outputs()
        arr = (C.c_void_p * max(1, len(outs)))(*[o.data_ptr() for o in outs])
        rid = C.c_int64()
        check(lib().dtcf_request_submit(self.server.h, self.h, arr, C.byref(rid)))
        return rid.value, outs

    def run(self, timeout_ms=600000.0):
        """Submit, wait, return {output name: host array}."""
        torch.cuda.synchronize()
        rid, outs = self.submit()
        check(self.server.wait(rid, timeout_ms))
        self.server.release(rid)
        return {n: o.cpu().numpy().reshape(s) for n, o, s in zip(self.outputs, outs, self.out_shapes)}

    def destroy(self):
        if self.h:
            lib().dtcf_program_destroy(self.h)
            self.h = None
