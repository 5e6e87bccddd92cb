"""Precision study: one CCD-like amplitude iteration in several precisions.

System (o occupied, v virtual spin orbitals): occupied eps ~ U[-2, -0.5],
virtual eps ~ U[0.2, 3]; integrals V_vvvv[a,b,c,d], V_vvoo[a,b,i,j] and
W[k,j,c,b] antisymmetrised N(0, 0.05^2); MP2 amplitudes
T[a,b,i,j] = V_vvoo / D.  One iteration:
    R[a,b,i,j] = V_vvoo + 1/2 sum_cd V[a,b,c,d] T[c,d,i,j]              (ladder)
                        + sum_kc T'[i,k,a,c] W[k,j,c,b]  -> R[a,b,i,j]  (ring)
    T_new = R / D,   E = 1/4 sum V_oovv[i,j,a,b] T_new[a,b,i,j]
where T' is T viewed as [i,k,a,c].  The two contractions run in the
precision under study; the denominator update and the energy always run in
FP64 on the device.  Errors are stated against the FP64 MFMA run of the same
inputs (max |dR|, max |dR| / max |R|, |dE|), in total and per term: each
contraction accumulates into its own buffer, and since E is linear in R the
energy splits as E = E(V_vvoo) + E(ladder) + E(ring), each E(X) =
1/4 sum V_oovv[i,j,a,b] X[a,b,i,j] / D_ijab.

Run: python -m dtcf_amd.precision --o 64 --v 256 --out study.json
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import mixed, ops
from ._lib import check, lib

VARIANTS = [
    ("fp64", None, 0, 0),
    ("fp32", mixed.F32, 1, 0),
    ("bf16x3", mixed.BF16X3, 3, 0),
    ("fp16x3", mixed.F16X3, 3, 0),
    ("int8x4", mixed.I8, 4, 0),      # 10 products, fused
    ("int8x4/L6", mixed.I8, 4, 6),   # 13 products, fused
    ("int8x4/L6, ring L5", mixed.I8, 4, (6, 5)),  # ladder 13 products, ring 10
    ("int8x5/L5", mixed.I8, 5, 5),   # 10 products, fused
    ("int8x5", mixed.I8, 5, 0),      # 15 products, fused
    ("int8x6", mixed.I8, 6, 0),      # 21 products, one GEMM each
]


def _sync():
    torch.cuda.synchronize()


class System:
    def __init__(self, o, v, dev="cuda", seed=100):
        self.o, self.v = o, v
        self.eo = mixed.fill_uniform(o, seed, -2.0, -0.5, dev)
        self.ev = mixed.fill_uniform(v, seed + 1, 0.2, 3.0, dev)
        raw = mixed.fill_normal(v ** 4, seed + 2, 0.05, dev)
        self.Vvvvv = mixed.antisymmetrize4(raw, (v, v, v, v))
        del raw
        raw = mixed.fill_normal(v * v * o * o, seed + 3, 0.05, dev)
        self.Vvvoo = mixed.antisymmetrize4(raw, (v, v, o, o))
        raw = mixed.fill_normal(o * o * v * v, seed + 4, 0.05, dev)
        W = mixed.antisymmetrize4(raw, (o, o, v, v))  # W[k,j,c,b]
        del raw
        # The ring's integral is kept as Wr[b,j,c,k] = W[k,j,c,b]: integrals
        # do not change between iterations, so they are stored once in the
        # layout the ring GEMM reads (rows (b,j), K (c,k), K contiguous).
        self.Wr = ops.rrr_transpose(W, [3, 1, 2, 0], [o, o, v, v])
        del W
        # V_oovv[i,j,a,b] = V_vvoo[a,b,i,j]
        self.Voovv = ops.rrr_transpose(self.Vvvoo, [2, 3, 0, 1], [v, v, o, o])
        self.T = ops.denom_update(self.Vvvoo, self.eo, self.ev)       # [c,d,i,j]
        oo, vv = o * o, v * v
        # ladder GEMM: C (ab x ij) += A (ab x cd) * B (ij x cd)^T
        self.lad_a = ops.View(vv, vv, np.arange(vv) * vv, np.arange(vv))
        self.lad_b = ops.View(oo, vv, np.arange(oo), np.arange(vv) * oo)
        # ring GEMM R[a,b,i,j] += sum_{k,c} T[a,c,i,k] W[k,j,c,b]:
        # rows (a,i), cols (b,j), K (c,k), the inner index fastest in each, so
        # T's rows are K-contiguous in runs of o, Wr's rows are K-contiguous,
        # and a C tile covers whole o x o blocks of R (contiguous writes).
        #   T[a,c,i,k] at ((a*v + c)*o + i)*o + k
        ai = np.array([a * v * o * o + i * o for a in range(v) for i in range(o)])
        ck = np.array([c * o * o + k for c in range(v) for k in range(o)])
        #   R[a,b,i,j] at ((a*v + b)*o + i)*o + j
        bj_r = np.array([b * o * o + j for b in range(v) for j in range(o)])
        n = o * v
        self.ring_a = ops.View(n, n, ai, ck)
        self.ring_b = ops.View(n, n, np.arange(n) * n, np.arange(n))   # Wr: dense [(b,j)][(c,k)]
        self.ring_c = ops.View(n, n, ai, bj_r)
        self.lad_flops = 2.0 * vv * oo * vv
        self.ring_flops = 2.0 * n * n * n
        _sync()

    def iterate(self, prec, nslice, level, terms=("ladder", "ring")):
        """Returns (R, T_new, E, seconds per phase, {term: its contribution
        to R}).  Phases are timed with a device synchronisation on each side,
        so they add up to the contraction time.  level: the largest digit
        level kept (0: nslice + 1), or a pair (ladder, ring)."""
        o, v = self.o, self.v
        lev_lad, lev_ring = level if isinstance(level, tuple) else (level, level)
        parts = {t: torch.zeros_like(self.Vvvoo) for t in terms}
        ph = {}

        def timed(name, fn):
            _sync()
            t = time.perf_counter()
            r = fn()
            _sync()
            ph[name] = ph.get(name, 0.0) + time.perf_counter() - t
            return r

        if prec is None:
            if "ladder" in parts:
                timed("gemm_ladder", lambda: ops.gemm_view(v * v, o * o, v * v, 0.5, self.Vvvvv,
                                                           self.lad_a, self.T, self.lad_b,
                                                           parts["ladder"], None, ldc=o * o))
            if "ring" in parts:
                timed("gemm_ring", lambda: ops.gemm_view(o * v, o * v, o * v, 1.0, self.T,
                                                         self.ring_a, self.Wr, self.ring_b,
                                                         parts["ring"], self.ring_c))
        else:
            # integrals (V_vvvv, W) are the same in every CC iteration: their
            # slices can stay resident; the amplitudes are split every time
            if "ladder" in parts:
                sa = timed("split_integrals", lambda: mixed.Split(prec, self.Vvvvv, self.lad_a, nslice))
                sb = timed("split_amplitudes", lambda: mixed.Split(prec, self.T, self.lad_b, nslice))
                timed("gemm_ladder", lambda: mixed.gemm_split(sa, sb, parts["ladder"], alpha=0.5,
                                                              ldc=o * o, max_level=lev_lad))
                del sa, sb
            if "ring" in parts:
                ra = timed("split_amplitudes", lambda: mixed.Split(prec, self.T, self.ring_a, nslice))
                rb = timed("split_integrals", lambda: mixed.Split(prec, self.Wr, self.ring_b, nslice))
                timed("gemm_ring", lambda: mixed.gemm_split(ra, rb, parts["ring"], alpha=1.0,
                                                            cview=self.ring_c, max_level=lev_ring))
                del ra, rb
        R = self.Vvvoo.clone()
        for x in parts.values():
            R += x
        Tn = ops.denom_update(R, self.eo, self.ev)
        E = ops.energy(self.Voovv, Tn, o, v).item()
        return R, Tn, E, ph, parts

    def energy_of(self, X):
        """1/4 sum V_oovv[i,j,a,b] X[a,b,i,j] / D_ijab: the energy carried by
        one contribution X to the residual (E is linear in R)."""
        return ops.energy(self.Voovv, ops.denom_update(X, self.eo, self.ev), self.o, self.v).item()


def study(o, v, variants=VARIANTS, reps=1, log=print):
    sysm = System(o, v)
    flops = sysm.lad_flops + sysm.ring_flops
    out = {"o": o, "v": v, "flops_ladder": sysm.lad_flops, "flops_ring": sysm.ring_flops,
           "variants": []}
    R64 = E64 = t64 = None
    P64 = {}
    for name, prec, ns, lev in variants:
        best = None
        for _ in range(max(1, reps)):
            R, Tn, E, ph, parts = sysm.iterate(prec, ns, lev)
            tc = sum(ph.values())
            if best is None or tc < best[0]:
                best = (tc, ph)
        tc, ph = best
        tcached = tc - ph.get("split_integrals", 0.0)
        rec = dict(name=name, seconds=tc, seconds_integrals_resident=tcached,
                   phases={k: round(x, 5) for k, x in ph.items()},
                   tflops=flops / tc / 1e12, energy=E)
        if prec is None:
            R64, E64, t64 = R, E, tc
            P64 = {k: (x, sysm.energy_of(x)) for k, x in parts.items()}
            rec.update(max_abs_err=0.0, max_rel_err=0.0, energy_err=0.0, speedup=1.0,
                       speedup_integrals_resident=1.0, products=1)
            rec["terms"] = {k: {"energy": e} for k, (x, e) in P64.items()}
        else:
            d = (R - R64).abs().max().item()
            rec.update(max_abs_err=d, max_rel_err=d / R64.abs().max().item(),
                       energy_err=abs(E - E64), speedup=t64 / tc,
                       speedup_integrals_resident=t64 / tcached,
                       products=(mixed.n_products(prec, ns, lev) if not isinstance(lev, tuple)
                                 else [mixed.n_products(prec, ns, x) for x in lev]))
            rec["terms"] = {}
            for k, x in parts.items():
                x64, e64 = P64[k]
                dk = (x - x64).abs().max().item()
                rec["terms"][k] = {"max_rel_err": dk / x64.abs().max().item(),
                                   "energy_err": abs(sysm.energy_of(x) - e64)}
        out["variants"].append(rec)
        log(json.dumps(rec))
        del R, Tn, parts
        torch.cuda.empty_cache()
    return out


def reductions(o, v, reps=5, log=print):
    """The FP64 streaming kernels of an iteration at [v,v,o,o] arrays:
    denominator update (R read, T written: 16 B per element), energy
    (V_oovv and T read: 16 B) and the residual norm (R read: 8 B); each
    timed with events over `reps` calls.  Returns GB/s per kernel."""
    n = v * v * o * o
    dev = "cuda"
    R = mixed.fill_normal(n, 1, 0.05, dev)
    Vm = mixed.fill_normal(n, 2, 0.05, dev)
    eo = mixed.fill_uniform(o, 3, -2.0, -0.5, dev)
    ev = mixed.fill_uniform(v, 4, 0.2, 3.0, dev)
    T = torch.empty_like(R)
    out = torch.zeros(1, dtype=torch.float64, device=dev)
    res = {}
    for name, fn, nbytes in (("denom_update", lambda: ops.denom_update(R, eo, ev, T), 16.0 * n),
                             ("energy", lambda: ops.energy(Vm, T, o, v, out), 16.0 * n),
                             ("norm2", lambda: ops.norm2(R, out), 8.0 * n)):
        fn()
        _sync()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        _sync()
        ms = e0.elapsed_time(e1) / reps
        res[name] = {"ms": ms, "GB/s": nbytes / ms / 1e6, "bytes": nbytes}
        log(f"{name}: {ms:.3f} ms, {nbytes / ms / 1e6:.0f} GB/s ({nbytes / 1e9:.2f} GB)")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--o", type=int, default=64)
    ap.add_argument("--v", type=int, default=256)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--reductions", action="store_true",
                    help="time the denominator update, energy and norm kernels only")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    check(lib().dtcf_init(0))
    if args.reductions:
        reductions(args.o, args.v, max(args.reps, 5), log=lambda s: print(s, flush=True))
        return 0
    res = study(args.o, args.v, reps=args.reps, log=lambda s: print(s, flush=True))
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
