"""Block-sparse tensors under an abelian point group (D2h) and the grouped
launch of a term's block pairs.

Orbitals of each space (occupied ``o``, virtual ``v``) are split into 8
irreps with seeded counts; irrep labels are 3-bit codes whose product is
XOR, so a block of a tensor with irreps (g1, ..., gd) is symmetry-allowed iff
g1 ^ ... ^ gd == 0 (about 1/8 of the blocks).  Only allowed blocks are
stored: contiguously, in lexicographic order of their irrep tuples, each a
dense row-major array over its index ranges.

``BlockTerm`` turns ``C[c_idx] += alpha * A[a_idx] * B[b_idx]`` into the op
and pair lists of ``dtcf_grouped_plan``: for every allowed C block, every
assignment of irreps to the contracted indices that makes both the A block
and the B block allowed gives one pair.  The GEMM axes are the A-external
indices (rows, in C's order), the B-external indices (columns) and the
contracted indices (in A's order); each block's strides are folded into the
descriptors, so no block is ever transposed in memory.
"""
from __future__ import annotations

import ctypes as C
import itertools
import json
import os

import numpy as np
import torch

from . import _lib
from ._lib import check, lib

NIRREP = 8


class Axis(C.Structure):
    _fields_ = [("nd", C.c_int32), ("ext", C.c_int32 * 3), ("stride", C.c_int64 * 3)]


class BlockOp(C.Structure):
    _fields_ = [("base", C.c_int64), ("rows", Axis), ("cols", Axis)]


class BlockPair(C.Structure):
    _fields_ = [("c_op", C.c_int32), ("a_op", C.c_int32), ("b_op", C.c_int32)]


class GemmProblem(C.Structure):
    _fields_ = [("m", C.c_int64), ("n", C.c_int64), ("k", C.c_int64), ("A", C.c_void_p),
                ("lda", C.c_int64), ("B", C.c_void_p), ("ldb", C.c_int64), ("C", C.c_void_p),
                ("ldc", C.c_int64)]


def _declare():
    L = lib()
    vp = C.c_void_p
    L.dtcf_grouped_plan.argtypes = [vp, C.c_int64, vp, C.c_int64, C.c_int, C.POINTER(vp)]
    L.dtcf_grouped_info.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int64)]
    L.dtcf_grouped_run.argtypes = [vp, C.c_double, vp, vp, vp, vp]
    L.dtcf_grouped_destroy.argtypes = [vp]
    L.dtcf_blockmove_plan.argtypes = [vp, vp, vp, C.c_int64, C.POINTER(vp)]
    L.dtcf_blockmove_info.argtypes = [vp, C.POINTER(C.c_int64)]
    L.dtcf_blockmove_gather_f64.argtypes = [vp, vp, vp, vp]
    L.dtcf_blockmove_scatter_f64.argtypes = [vp, C.c_double, vp, vp, vp]
    L.dtcf_blockmove_destroy.argtypes = [vp]
    L.dtcf_gemm_group_f64.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.c_int, vp]
    L.dtcf_gemm_group_ex_f64.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                         C.c_int, vp]
    for f in ("dtcf_gemm_group_f64", "dtcf_gemm_group_ex_f64", "dtcf_grouped_plan", "dtcf_grouped_info", "dtcf_grouped_run", "dtcf_grouped_destroy",
              "dtcf_blockmove_plan", "dtcf_blockmove_info", "dtcf_blockmove_gather_f64",
              "dtcf_blockmove_scatter_f64", "dtcf_blockmove_destroy"):
        getattr(L, f).restype = C.c_int


_declare()


def irrep_counts(total: int, seed: int, lo: int, hi: int):
    """Seeded per-irrep orbital counts in [lo, hi] summing to total."""
    rng = np.random.default_rng(seed)
    while True:
        c = rng.integers(lo, hi + 1, size=NIRREP)
        diff = total - int(c.sum())
        i = 0
        while diff != 0 and i < 10000:
            j = rng.integers(NIRREP)
            step = 1 if diff > 0 else -1
            if lo <= c[j] + step <= hi:
                c[j] += step
                diff -= step
            i += 1
        if diff == 0:
            return [int(x) for x in c]


class Space:
    def __init__(self, occ_counts, vir_counts):
        self.counts = {"o": list(occ_counts), "v": list(vir_counts)}

    @property
    def o(self):
        return sum(self.counts["o"])

    @property
    def v(self):
        return sum(self.counts["v"])

    @staticmethod
    def seeded(o=64, v=256, seed=5):
        return Space(irrep_counts(o, seed, max(1, o // 16), max(2, o // 5)),
                     irrep_counts(v, seed + 1, max(1, v // 16), max(2, v // 5)))


class BSTensor:
    """Allowed blocks of a tensor whose dims have kinds in {'o','v'}."""

    def __init__(self, space: Space, kinds: str, device="cuda", dtype=torch.float64):
        self.space = space
        self.kinds = kinds
        self.blocks = []       # irrep tuples
        self.offset = {}       # irreps -> element offset
        self.shape = {}        # irreps -> extents
        off = 0
        for g in itertools.product(range(NIRREP), repeat=len(kinds)):
            x = 0
            for gi in g:
                x ^= gi
            if x != 0:
                continue
            shp = tuple(space.counts[k][gi] for k, gi in zip(kinds, g))
            self.blocks.append(g)
            self.offset[g] = off
            self.shape[g] = shp
            off += int(np.prod(shp))
        self.size = off
        self.data = torch.zeros(off, dtype=dtype, device=device) if device else None

    def fill_normal(self, seed, sigma=0.05):
        from .synth import normal
        self.data.copy_(torch.from_numpy(normal(self.size, seed, sigma)))
        return self

    def block(self, g, host=None):
        buf = self.data.cpu().numpy() if host is None else host
        o = self.offset[g]
        return buf[o:o + int(np.prod(self.shape[g]))].reshape(self.shape[g])


def _axis(exts, strides):
    a = Axis()
    a.nd = len(exts)
    for i, (e, s) in enumerate(zip(exts, strides)):
        a.ext[i] = int(e)
        a.stride[i] = int(s)
    return a


def _strides(shape):
    st = [1] * len(shape)
    for d in range(len(shape) - 2, -1, -1):
        st[d] = st[d + 1] * shape[d + 1]
    return st


def block_op(t: BSTensor, idx, g, row_names, col_names):
    """Descriptor of block g of t (index names idx) as a rows x cols matrix:
    rows over row_names, columns over col_names (each in the given order)."""
    shp = t.shape[g]
    st = _strides(shp)
    pos = {x: i for i, x in enumerate(idx)}
    op = BlockOp()
    op.base = t.offset[g]
    op.rows = _axis([shp[pos[x]] for x in row_names], [st[pos[x]] for x in row_names])
    op.cols = _axis([shp[pos[x]] for x in col_names], [st[pos[x]] for x in col_names])
    return op


class BlockMove:
    """dtcf_blockmove plan: blocks of one tensor <-> placed regions of a
    dense buffer."""

    def __init__(self, ops, bases, pitches):
        n = len(ops)
        self.ops = (BlockOp * max(1, n))(*ops)
        self.bases = (C.c_int64 * max(1, n))(*bases)
        self.pitches = (C.c_int64 * max(1, n))(*pitches)
        h = C.c_void_p()
        check(lib().dtcf_blockmove_plan(C.cast(self.ops, C.c_void_p), C.cast(self.bases, C.c_void_p),
                                        C.cast(self.pitches, C.c_void_p), n, C.byref(h)))
        self.h = h
        sp = C.c_int64()
        check(lib().dtcf_blockmove_info(h, C.byref(sp)))
        self.span = sp.value

    def gather(self, sparse, dense):
        check(lib().dtcf_blockmove_gather_f64(self.h, C.c_void_p(sparse.data_ptr()),
                                              C.c_void_p(dense.data_ptr()), _s()))

    def scatter(self, alpha, dense, sparse):
        check(lib().dtcf_blockmove_scatter_f64(self.h, float(alpha), C.c_void_p(dense.data_ptr()),
                                               C.c_void_p(sparse.data_ptr()), _s()))

    def __del__(self):
        try:
            if self.h:
                lib().dtcf_blockmove_destroy(self.h)
                self.h = None
        except Exception:
            pass


def _s():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _xor(g):
    x = 0
    for v in g:
        x ^= v
    return x


class IrrepDense:
    """A block-sparse term as one dense GEMM per irrep.

    With h the product (XOR) of the irreps of C's row indices, an allowed C
    block has column irreps of product h too, and an A (B) block feeding it
    has contracted irreps of product h.  For each h every (row block, column
    block, contracted block) combination occurs, so the term is
        C_h[M_h x N_h] += A_h[M_h x K_h] * B_h[N_h x K_h]^T
    over the blocks of class h laid side by side.  A and B are gathered into
    the A_h / B_h matrices (one launch each), one grouped launch computes
    every C_h (FP64 MFMA, overwriting), and C_h is added back into C's blocks
    (one launch).  An operand whose fastest (last) index is external is
    gathered as its transpose, K_h x M_h (or K_h x N_h), so the gather runs
    along the blocks' contiguous runs, and the GEMM reads it in that layout
    (a B gather through transposed tiles ran 1.8-2.0 TB/s).  The counted
    flops are exactly those of the block pairs; the GEMMs also run over K_h's
    zero padding to a multiple of 16 (< 1% extra at these sizes)."""

    def __init__(self, term):
        t = term
        sp, kind = t.C.space, t.kind
        a_ext, b_ext, contr = t.a_ext, t.b_ext, t.contr

        def tuples(names, h):
            return [g for g in itertools.product(range(NIRREP), repeat=len(names)) if _xor(g) == h]

        def size(names, g):
            return int(np.prod([sp.counts[kind[x]][gi] for x, gi in zip(names, g)]))

        def assemble(idx, names1, g1, names2, g2):
            d = dict(zip(names1, g1))
            d.update(zip(names2, g2))
            return tuple(d[x] for x in idx)

        aops, abase, apitch = [], [], []
        bops, bbase, bpitch = [], [], []
        cops, cbase, cpitch = [], [], []
        # operands whose fastest index is external: gathered transposed
        self.transa = int(t.a_idx[-1] in a_ext)
        self.transb = int(t.b_idx[-1] in b_ext)
        self.gemms = []  # (m, n, padded k, a_off, b_off, c_off, lda, ldb)
        self.flops = 0.0
        la = lb = lc = 0
        for h in range(NIRREP):
            R, Cc, Kc = tuples(a_ext, h), tuples(b_ext, h), tuples(contr, h)
            rs = [size(a_ext, g) for g in R]
            cs = [size(b_ext, g) for g in Cc]
            ks = [size(contr, g) for g in Kc]
            M, N, K = sum(rs), sum(cs), sum(ks)
            if M == 0 or N == 0 or K == 0:
                continue
            # K padded to the GEMM's K step with zero columns (never written
            # by the gathers), so every tile takes the unchecked 16-byte loads
            Kp = -(-K // 16) * 16
            # transposed layouts: K_h rows (padding rows zero) of an even pitch
            lda = M + M % 2 if self.transa else Kp
            ldb = N + N % 2 if self.transb else Kp
            roff = np.concatenate([[0], np.cumsum(rs)])
            coff = np.concatenate([[0], np.cumsum(cs)])
            koff = np.concatenate([[0], np.cumsum(ks)])
            for i, gr in enumerate(R):
                for k, gk in enumerate(Kc):
                    ga = assemble(t.a_idx, a_ext, gr, contr, gk)
                    if self.transa:
                        aops.append(block_op(t.A, t.a_idx, ga, contr, a_ext))
                        abase.append(la + int(koff[k]) * lda + int(roff[i]))
                    else:
                        aops.append(block_op(t.A, t.a_idx, ga, a_ext, contr))
                        abase.append(la + int(roff[i]) * lda + int(koff[k]))
                    apitch.append(lda)
            for j, gc in enumerate(Cc):
                for k, gk in enumerate(Kc):
                    gb = assemble(t.b_idx, b_ext, gc, contr, gk)
                    if self.transb:
                        bops.append(block_op(t.B, t.b_idx, gb, contr, b_ext))
                        bbase.append(lb + int(koff[k]) * ldb + int(coff[j]))
                    else:
                        bops.append(block_op(t.B, t.b_idx, gb, b_ext, contr))
                        bbase.append(lb + int(coff[j]) * ldb + int(koff[k]))
                    bpitch.append(ldb)
            for i, gr in enumerate(R):
                for j, gc in enumerate(Cc):
                    gcb = assemble(t.c_idx, a_ext, gr, b_ext, gc)
                    cops.append(block_op(t.C, t.c_idx, gcb, a_ext, b_ext))
                    cbase.append(lc + int(roff[i]) * N + int(coff[j]))
                    cpitch.append(N)
            self.gemms.append((M, N, Kp, la, lb, lc, lda, ldb))
            self.flops += 2.0 * M * N * K
            la += (Kp if self.transa else M) * lda
            lb += (Kp if self.transb else N) * ldb
            lc += M * N
        self.mA, self.mB, self.mC = (BlockMove(aops, abase, apitch), BlockMove(bops, bbase, bpitch),
                                     BlockMove(cops, cbase, cpitch))
        # every move stays inside the dense buffers allocated below
        assert self.mA.span <= max(1, la) and self.mB.span <= max(1, lb) and self.mC.span <= max(1, lc)
        dev = t.C.data.device
        self.dA = torch.zeros(max(1, la), dtype=torch.float64, device=dev)
        self.dB = torch.zeros(max(1, lb), dtype=torch.float64, device=dev)
        self.dC = torch.empty(max(1, lc), dtype=torch.float64, device=dev)
        a, b, c = self.dA.data_ptr(), self.dB.data_ptr(), self.dC.data_ptr()
        self.probs = (GemmProblem * len(self.gemms))(*[
            GemmProblem(m, n, k, a + 8 * ao, lda, b + 8 * bo, ldb, c + 8 * co, n)
            for m, n, k, ao, bo, co, lda, ldb in self.gemms])

    # C tiles of the grouped launch tried on the first run (128x128, 64x64):
    # the wider tile reuses more per load, the narrower one fills the last
    # round of tiles better, and which wins depends on the irrep sizes
    GEMM_CFGS = (0, 3)

    def _gemms(self, cfg):
        # op(A_h) = A_h or (K x M)^T; op(B_h) = B_h^T or the K x N gather itself
        check(lib().dtcf_gemm_group_ex_f64(C.cast(self.probs, C.c_void_p), len(self.gemms),
                                           self.transa, 1 - self.transb, 1.0, 1, cfg, _s()))

    def _tune(self):
        """Times each candidate tile on the gathered operands (C_h is
        overwritten, so this has no effect on the result) and keeps the
        fastest."""
        best = None
        for cfg in self.GEMM_CFGS:
            self._gemms(cfg)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            self._gemms(cfg)
            self._gemms(cfg)
            e1.record()
            e1.synchronize()
            ms = e0.elapsed_time(e1) / 2
            if best is None or ms < best[0]:
                best = (ms, cfg)
        self.gemm_cfg = best[1]

    def set_precision(self, prec=0, nslice=4, max_level=0):
        """prec 0: FP64 (one grouped launch).  A mixed mode (mixed.F32,
        BF16X3, F16X3, I8 with nslice digits and products up to max_level):
        every A_h and B_h is split into slices (one two-pass split per
        matrix, dense row-major views) and each C_h is the sum of the slice
        products (dtcf_gemm_split_f64, FP64 recombination).  The slice
        buffers are allocated here, once."""
        from . import mixed, ops
        self.prec, self.max_level = prec, max_level
        self._mx = None
        if prec == 0:
            return
        ns = nslice if prec == mixed.I8 else (1 if prec == mixed.F32 else 3)
        self.nslice = ns
        dev = self.dA.device
        mx = []
        for m, n, k, ao, bo, co, lda, ldb in self.gemms:
            ent = {}
            for tag, rows, base, ld, tr in (("a", m, ao, lda, self.transa),
                                            ("b", n, bo, ldb, self.transb)):
                r, c = np.arange(rows, dtype=np.int64), np.arange(k, dtype=np.int64)
                view = (ops.View(rows, k, base + r, c * ld) if tr else
                        ops.View(rows, k, base + r * ld, c))
                nb = lib().dtcf_split_bytes(prec, ns, rows, k)
                buf = torch.empty(nb, dtype=torch.uint8, device=dev)
                rexp = torch.zeros((rows + 127) // 128 * 128, dtype=torch.int32, device=dev)
                ent[tag] = (view, buf, rexp)
            mx.append(ent)
        self._mx = mx

    def _split(self, tag, src):
        for ent in self._mx:
            view, buf, rexp = ent[tag]
            check(lib().dtcf_split_f64(self.prec, self.nslice, C.c_void_p(src.data_ptr()), view.h,
                                       C.c_void_p(buf.data_ptr()), C.c_void_p(rexp.data_ptr()), _s()))

    def _mixed_gemms(self):
        self.dC.zero_()  # the slice products accumulate
        c0 = self.dC.data_ptr()
        for (m, n, k, ao, bo, co, _, _), ent in zip(self.gemms, self._mx):
            _, ab, ae = ent["a"]
            _, bb, be = ent["b"]
            check(lib().dtcf_gemm_split_f64(self.prec, self.nslice, m, n, k, 1.0,
                                            C.c_void_p(ab.data_ptr()), C.c_void_p(ae.data_ptr()),
                                            C.c_void_p(bb.data_ptr()), C.c_void_p(be.data_ptr()),
                                            int(self.max_level), C.c_void_p(c0 + 8 * co), None, n,
                                            _s()))

    def run(self, alpha, A, B, Cb, a_resident=False):
        """Gathers A and B into the per-irrep matrices, computes every
        C_h = A_h B_h^T in one grouped launch (overwriting C_h), then adds
        alpha * C_h into C's blocks, all on the caller's stream.
        a_resident: the caller promises A is the buffer of the previous run,
        unchanged since (fixed integrals), so its gathered form (and, in a
        mixed mode, its slices) is reused."""
        mixed_mode = getattr(self, "_mx", None) is not None
        key = (A.data_ptr(), getattr(self, "prec", 0), getattr(self, "nslice", 0))
        if not (a_resident and getattr(self, "_a_src", None) == key):
            self.mA.gather(A, self.dA)
            if mixed_mode:
                self._split("a", self.dA)
            self._a_src = key
        self.mB.gather(B, self.dB)
        if mixed_mode:
            self._split("b", self.dB)
            self._mixed_gemms()
        else:
            if getattr(self, "gemm_cfg", None) is None:
                self._tune()
            self._gemms(self.gemm_cfg)
        self.mC.scatter(alpha, self.dC, Cb)


class BlockTerm:
    """C[c_idx] += alpha * A[a_idx] * B[b_idx] over allowed blocks."""

    def __init__(self, Ct: BSTensor, c_idx: str, At: BSTensor, a_idx: str, Bt: BSTensor,
                 b_idx: str, cfg: int | None = None):
        assert len(c_idx) == len(Ct.kinds) and len(a_idx) == len(At.kinds) and len(b_idx) == len(Bt.kinds)
        self.C, self.A, self.B = Ct, At, Bt
        self.c_idx, self.a_idx, self.b_idx = c_idx, a_idx, b_idx
        contr = [x for x in a_idx if x in b_idx]
        a_ext = [x for x in c_idx if x in a_idx]
        b_ext = [x for x in c_idx if x in b_idx]
        assert sorted(a_ext + b_ext) == sorted(c_idx), "C must hold the external indices"
        assert 1 <= len(a_ext) <= 3 and 1 <= len(b_ext) <= 3 and 1 <= len(contr) <= 3
        self.contr, self.a_ext, self.b_ext = contr, a_ext, b_ext
        kind = {}
        for idx, t in ((a_idx, At), (b_idx, Bt), (c_idx, Ct)):
            for x, k in zip(idx, t.kinds):
                assert kind.setdefault(x, k) == k, f"index {x} used with two kinds"
        self.kind = kind
        ops, pairs = [], []
        opid = {}

        def op_for(t: BSTensor, tag, idx, g, row_names, col_names):
            key = (tag, g, tuple(row_names), tuple(col_names))
            if key in opid:
                return opid[key]
            shp = t.shape[g]
            st = _strides(shp)
            pos = {x: i for i, x in enumerate(idx)}
            op = BlockOp()
            op.base = t.offset[g]
            op.rows = _axis([shp[pos[x]] for x in row_names], [st[pos[x]] for x in row_names])
            op.cols = _axis([shp[pos[x]] for x in col_names], [st[pos[x]] for x in col_names])
            opid[key] = len(ops)
            ops.append(op)
            return opid[key]

        self.flops = 0.0
        self.pair_flops = 0.0
        for gc in Ct.blocks:
            irr = dict(zip(c_idx, gc))
            cop = op_for(Ct, "C", c_idx, gc, a_ext, b_ext)
            for gk in itertools.product(range(NIRREP), repeat=len(contr)):
                irr.update(zip(contr, gk))
                ga = tuple(irr[x] for x in a_idx)
                gb = tuple(irr[x] for x in b_idx)
                if ga not in At.offset or gb not in Bt.offset:
                    continue
                aop = op_for(At, "A", a_idx, ga, a_ext, contr)
                bop = op_for(Bt, "B", b_idx, gb, b_ext, contr)
                ext = {**dict(zip(a_idx, At.shape[ga])), **dict(zip(b_idx, Bt.shape[gb]))}
                self.pair_flops += 2.0 * float(np.prod([ext[x] for x in ext]))
                p = BlockPair()
                p.c_op, p.a_op, p.b_op = cop, aop, bop
                pairs.append(p)
        self.ops = (BlockOp * len(ops))(*ops)
        self.pairs = (BlockPair * len(pairs))(*pairs)
        self.npairs = len(pairs)
        self.nops = len(ops)
        self.h = None
        self._dense = None
        self.cfg = pick_config(self) if cfg is None else cfg
        self._plan()

    def _plan(self):
        if self.cfg == 4:
            self.h = None
            self._dense = IrrepDense(self)
            self.nwork = len(self._dense.gemms)
            assert abs(self._dense.flops - self.pair_flops) <= 1e-6 * max(1.0, self.pair_flops)
            self.flops = self._dense.flops
            return
        self._dense = None
        h = C.c_void_p()
        check(lib().dtcf_grouped_plan(C.cast(self.ops, C.c_void_p), self.nops,
                                      C.cast(self.pairs, C.c_void_p), self.npairs, int(self.cfg),
                                      C.byref(h)))
        self.h = h
        fl = C.c_double()
        nw = C.c_int64()
        spans = (C.c_int64 * 3)()
        check(lib().dtcf_grouped_info(h, C.byref(fl), C.byref(nw), spans))
        self.flops = fl.value
        self.nwork = nw.value
        assert spans[0] <= self.A.size and spans[1] <= self.B.size and spans[2] <= self.C.size

    def set_precision(self, prec=0, nslice=4, max_level=0):
        """Mixed-precision contraction (see IrrepDense.set_precision); only
        the dense-by-irrep configuration (4) has one."""
        if prec != 0 and self.cfg != 4:
            raise ValueError("block-sparse mixed precision runs in the dense-by-irrep "
                             "configuration (cfg 4): call replan(4) first")
        if self._dense is not None:
            self._dense.set_precision(prec, nslice, max_level)

    def replan(self, cfg):
        if self.h:
            lib().dtcf_grouped_destroy(self.h)
            self.h = None
        self._dense = None
        self.cfg = cfg
        self._plan()

    def run(self, alpha=1.0, A=None, B=None, Cb=None, a_resident=False):
        """a_resident (dense-by-irrep configuration only): A is unchanged
        since the previous run (fixed integrals), so its gather is skipped."""
        A = self.A.data if A is None else A
        B = self.B.data if B is None else B
        Cb = self.C.data if Cb is None else Cb
        if self.cfg == 4:
            self._dense.run(alpha, A, B, Cb, a_resident)
            return
        check(lib().dtcf_grouped_run(self.h, float(alpha), C.c_void_p(A.data_ptr()),
                                     C.c_void_p(B.data_ptr()), C.c_void_p(Cb.data_ptr()),
                                     C.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def __del__(self):
        try:
            if self.h:
                lib().dtcf_grouped_destroy(self.h)
                self.h = None
        except Exception:
            pass


# ---------------------------------------------------------------- autotuner
# tile configurations of dtcf_grouped_plan
CONFIGS = {0: "64x64", 1: "128x128", 2: "64x64 packed", 3: "128x128 packed",
           4: "dense GEMM per irrep"}
_TUNE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_grouped.json")
_TUNED = None


def _bucket(term):
    """Size bucket of a term: log2 of the mean C-block rows, cols and K."""
    ops = term.ops
    rows = cols = ks = 0.0
    for p in term.pairs[:min(term.npairs, 4096)]:
        c, a = ops[p.c_op], ops[p.a_op]
        r = np.prod([c.rows.ext[i] for i in range(c.rows.nd)])
        cc = np.prod([c.cols.ext[i] for i in range(c.cols.nd)])
        k = np.prod([a.cols.ext[i] for i in range(a.cols.nd)])
        rows += r
        cols += cc
        ks += k
    n = max(1, min(term.npairs, 4096))
    return f"{int(np.log2(max(rows / n, 1)))}-{int(np.log2(max(cols / n, 1)))}-{int(np.log2(max(ks / n, 1)))}"


def pick_config(term):
    global _TUNED
    if _TUNED is None:
        try:
            with open(_TUNE_PATH) as f:
                _TUNED = json.load(f)
        except (OSError, ValueError):
            _TUNED = {}
    b = _bucket(term)
    if b in _TUNED:
        return int(_TUNED[b])
    # default: packed operands; large C blocks amortise better over 128x128 tiles
    r, c, _ = (int(x) for x in b.split("-"))
    return 3 if r >= 8 and c >= 8 else 2


def autotune(term, reps=3):
    """Time every tile configuration on the device; record the best for the
    term's size bucket in tuned_grouped.json and return (cfg, times)."""
    global _TUNED
    times = {}
    for cfg in CONFIGS:
        term.replan(cfg)
        term.run(0.0)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            term.run(0.0)
        e1.record()
        torch.cuda.synchronize()
        times[cfg] = e0.elapsed_time(e1) / reps
    best = min(times, key=times.get)
    term.replan(best)
    if _TUNED is None:
        _TUNED = {}
    _TUNED[_bucket(term)] = best
    return best, times


def save_tuning(path=None):
    with open(path or _TUNE_PATH, "w") as f:
        json.dump(_TUNED or {}, f, indent=1, sort_keys=True)


# ---------------------------------------------------------------- benchmark
BENCH_TERMS = {"ring ijab+=ikac*kjcb": (("oovv", "ijab"), ("oovv", "ikac"), ("oovv", "kjcb")),
               "ladder abij+=abcd*cdij": (("vvoo", "abij"), ("vvvv", "abcd"), ("vvoo", "cdij"))}


def bench_terms(dev="cuda", o=64, v=256, seed=5, steps=5):
    """The D2h ring and ladder terms at o, v: autotune each, then time
    `steps` runs (HIP events).  Returns per-term block pairs, performed flops,
    ms per run (pack launches included) and TFLOP/s."""
    from . import mixed
    sp = Space.seeded(o=o, v=v, seed=seed)
    out = {}
    for name, ((ck, ci), (ak, ai), (bk, bi)) in BENCH_TERMS.items():
        ts = []
        for sd, kinds in ((1, ak), (2, bk), (3, ck)):
            t = BSTensor(sp, kinds, dev)
            t.data.copy_(mixed.fill_normal(t.size, sd, 0.05))
            ts.append(t)
        A, B, Ct = ts
        term = BlockTerm(Ct, ci, A, ai, B, bi)
        cfg, times = autotune(term)
        term.run(1.0)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            term.run(1.0)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        out[name] = {"block_pairs": term.npairs, "flops_performed": term.flops, "ms": round(ms, 4),
                     "tflops": round(term.flops / ms / 1e9, 3), "tile_config": CONFIGS[cfg],
                     "autotune_ms": {CONFIGS[k]: round(x, 4) for k, x in times.items()}}
        if cfg == 4:
            # A (the integrals of the ladder term) gathered once and kept
            e0.record()
            for _ in range(steps):
                term.run(1.0, a_resident=True)
            e1.record()
            torch.cuda.synchronize()
            msr = e0.elapsed_time(e1) / steps
            out[name].update({"gemm_tile": {0: "128x128", 3: "64x64"}[term._dense.gemm_cfg],
                              "ms_a_resident": round(msr, 4),
                              "tflops_a_resident": round(term.flops / msr / 1e9, 3)})
        del A, B, Ct, term, ts
        torch.cuda.empty_cache()
    return out


# mixed modes of the dense-by-irrep configuration: (precision, digits, level)
MIXED_MODES = {"int8x5/L6": (4, 5, 6), "int8x4/L5": (4, 4, 5), "int8x3/L4": (4, 3, 4)}


def bench_mixed_terms(dev="cuda", o=64, v=256, seed=5, steps=5, modes=("int8x4/L5", "int8x3/L4")):
    """The D2h ring and ladder terms at o, v in the dense-by-irrep
    configuration: FP64 against each mixed mode.  Per term and mode: ms per
    run (gathers, splits, slice products and scatter), ms with A (the
    integrals) resident as gathered slices, speedup over the FP64 form of
    the same configuration and over the autotuned FP64 configuration, and
    the error of C against FP64 (largest element and Frobenius norm, both
    relative to C's largest element / norm)."""
    from . import mixed
    sp = Space.seeded(o=o, v=v, seed=seed)
    out = {}

    def timed(term, Cb, res, n):
        term.run(1.0, Cb=Cb, a_resident=res)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(n):
            term.run(1.0, Cb=Cb, a_resident=res)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    for name, ((ck, ci), (ak, ai), (bk, bi)) in BENCH_TERMS.items():
        A = BSTensor(sp, ak, dev)
        A.data.copy_(mixed.fill_normal(A.size, 1, 0.05))
        B = BSTensor(sp, bk, dev)
        B.data.copy_(mixed.fill_normal(B.size, 2, 0.05))
        Ct = BSTensor(sp, ck, dev)
        term = BlockTerm(Ct, ci, A, ai, B, bi)
        best_cfg, times = autotune(term)
        ms_best = times[best_cfg]
        term.replan(4)
        scratch = torch.zeros_like(Ct.data)
        ms64 = timed(term, scratch, False, steps)
        ms64r = timed(term, scratch, True, steps)
        Ct.data.zero_()
        term.run(1.0)
        ref = Ct.data.clone()
        rmax = float(ref.abs().max())
        rnorm = float(torch.linalg.vector_norm(ref))
        row = {"block_pairs": term.npairs, "flops": term.flops,
               "fp64_best": {"tile_config": CONFIGS[best_cfg], "ms": round(ms_best, 4),
                             "tflops": round(term.flops / ms_best / 1e9, 3)},
               "fp64_dense_by_irrep": {"ms": round(ms64, 4), "ms_a_resident": round(ms64r, 4)}}
        for mode in modes:
            prec, ns, lev = MIXED_MODES[mode]
            term.set_precision(prec, ns, lev)
            ms = timed(term, scratch, False, steps)
            msr = timed(term, scratch, True, steps)
            Ct.data.zero_()
            term.run(1.0)
            d = Ct.data - ref
            row[mode] = {"ms": round(ms, 4), "ms_a_resident": round(msr, 4),
                         "tflops_equiv": round(term.flops / ms / 1e9, 3),
                         "speedup_vs_fp64_dense_by_irrep": round(ms64 / ms, 3),
                         "speedup_a_resident": round(ms64r / msr, 3),
                         "speedup_vs_fp64_best": round(ms_best / ms, 3),
                         "rel_err_max": float(d.abs().max()) / rmax,
                         "rel_err_frobenius": float(torch.linalg.vector_norm(d)) / rnorm}
            term.set_precision(0)
        out[name] = row
        del A, B, Ct, term, scratch, ref
        torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="D2h block-sparse ring and ladder terms, grouped launch")
    ap.add_argument("--o", type=int, default=64)
    ap.add_argument("--v", type=int, default=256)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--mixed", nargs="*", default=None, metavar="MODE",
                    help=f"mixed-precision comparison in the dense-by-irrep form ({', '.join(MIXED_MODES)})")
    args = ap.parse_args()
    check(lib().dtcf_init(0))
    if args.mixed is not None:
        res = bench_mixed_terms("cuda", args.o, args.v, steps=args.steps,
                                modes=tuple(args.mixed) or ("int8x4/L5", "int8x3/L4"))
    else:
        res = bench_terms("cuda", args.o, args.v, steps=args.steps)
    print(json.dumps(res, indent=1), flush=True)
