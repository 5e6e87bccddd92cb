"""Thin FP64 products of the UHF CCSD iteration at o=96, v=384 (one side of
C 48 or 192 wide, short K), on the automatic kernel choice (the streaming
thin kernel) and on 64 x 64 tiles.  Inputs resident on the device; C
accumulated (C += A B), so C is read and written.  Reports per shape the
average time of each path, the rate against the algorithmic bytes (A, B
once, C read and written) and flops, and the largest difference between
the two paths relative to the largest result element.

Run: python -m dtcf_amd.thinbench [--reps 10] [--copy]
(--copy adds a plain 2 GiB device copy, for scaling HBM counters.)
"""
from __future__ import annotations

import argparse
import json

import torch

from . import ops
from ._lib import check, lib

# (m, n, k, transa, transb): rows x 48 x 192 and rows x 48 x 48 with A
# k-contiguous; 442368 x 192 x 48; 192 x 442368 x 48 (the long operand is B,
# C written along it)
SHAPES = ((7077888, 48, 192, 0, 0), (1769472, 48, 192, 0, 0), (1769472, 48, 48, 0, 0),
          (442368, 192, 48, 0, 0), (192, 442368, 48, 0, 0))


def run(shapes=SHAPES, reps=10, seed=5):
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    rows = []
    for (m, n, k, ta, tb) in shapes:
        A = torch.randn(m * k, dtype=torch.float64, device=dev, generator=g)
        B = torch.randn(k * n, dtype=torch.float64, device=dev, generator=g)
        lda, ldb = (m if ta else k), (k if tb else n)
        row = {"m": m, "n": n, "k": k, "transa": ta, "transb": tb}
        outs = {}
        for name, cfg in (("auto", -1), ("tiles_64x64", 3)):
            C = torch.zeros(m * n, dtype=torch.float64, device=dev)
            ops.set_gemm_config(cfg)
            try:
                ops.gemm(ta, tb, m, n, k, 1.0, A, lda, B, ldb, C, n)  # first call sets attributes
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(reps):
                    ops.gemm(ta, tb, m, n, k, 1.0, A, lda, B, ldb, C, n)
                e1.record()
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / reps
                C.zero_()
                ops.gemm(ta, tb, m, n, k, 1.0, A, lda, B, ldb, C, n)
                torch.cuda.synchronize()
            finally:
                ops.set_gemm_config(-1)
            outs[name] = C
            nbytes = 8 * (m * k + k * n + 2 * m * n)
            row[name] = {"ms": round(ms, 4), "TB_per_s": round(nbytes / ms / 1e9, 3),
                         "TFLOP_per_s": round(2 * m * n * k / ms / 1e9, 2)}
        ref = outs["tiles_64x64"]
        row["max_rel_diff"] = float((outs["auto"] - ref).abs().max() / ref.abs().max())
        row["speedup"] = round(row["tiles_64x64"]["ms"] / row["auto"]["ms"], 3)
        rows.append(row)
        del A, B, outs, ref, C
        torch.cuda.empty_cache()
    return {"workload": "thin FP64 products of the UHF CCSD iteration (o=96 v=384), C accumulated",
            "shapes": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--copy", action="store_true", help="also run a plain 2 GiB device copy")
    args = ap.parse_args()
    check(lib().dtcf_init(0))
    out = run(reps=args.reps)
    if args.copy:
        x = torch.empty(1 << 28, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        out["copy_calibration"] = "2 GiB read + 2 GiB written by one copy kernel"
    print(json.dumps(out))


if __name__ == "__main__":
    main()
