"""Served CCSD iteration stream: UHF T1 and T2 residual programs over several
synthetic systems, mixed in one queue with Poisson arrivals, under an HBM
budget.  Reports throughput (dense TFLOP/s, requests/s, statements/s) and
p50 / p99 latency per statement (ready -> done) and per request
(submit -> done).

Run: python -m dtcf_amd.servebench --rate 2.5 --requests 36
"""
from __future__ import annotations

import argparse
import json
import random
import sys
import time

import numpy as np
import torch

from . import mixed, serve, termlist
from ._lib import check, lib

SYSTEMS = ((32, 128), (64, 256), (96, 384))


def build(budget_gb=200.0, nslots=8, systems=SYSTEMS, precs=((0, 0),), log=print,
          resident_integrals=False):
    t1, t2 = termlist.load("t1"), termlist.load("t2")
    srv = serve.Server(int(budget_gb * (1 << 30)), nslots)
    progs = []
    for o, v in systems:
        t = time.perf_counter()
        sysm = serve.System(srv, o, v, [t1, t2], seed=o * 1000 + v)
        for pr in precs:
            prec, ns = pr[0], pr[1]
            lev = pr[2] if len(pr) > 2 else 0
            for which, tl in (("t1", t1), ("t2", t2)):
                res = serve.integral_inputs(tl) if resident_integrals and prec != 0 else ()
                p = serve.Program(srv, tl, sysm, prec=prec, nslice=ns, max_level=lev, resident=res)
                p.label = f"{which}/o{o}v{v}/" + ("fp64" if prec == 0 else
                                                   mixed.NAMES[prec] + f"x{ns}" + (f"L{lev}" if lev else ""))
                progs.append(p)
        log(f"system o={o} v={v}: inputs {sysm.bytes() / 1e9:.2f} GB, set up in "
            f"{time.perf_counter() - t:.2f} s; server resident {srv.usage()['resident'] / 1e9:.2f} GB")
    return srv, progs


def run_stream(srv, progs, nreq=36, rate=2.5, seed=1, log=print):
    """Submits nreq requests, each a uniformly chosen program, at Poisson
    arrival times of the given rate (per second; 0 = all at once)."""
    rng = random.Random(seed)
    picks = [rng.randrange(len(progs)) for _ in range(nreq)]
    gaps = [rng.expovariate(rate) if rate > 0 else 0.0 for _ in range(nreq)]
    torch.cuda.synchronize()
    subs = []
    t0 = time.perf_counter()
    due = 0.0
    for k, g in zip(picks, gaps):
        due += g
        dt = t0 + due - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        rid, outs = progs[k].submit()
        subs.append((k, rid, outs))
    for k, rid, outs in subs:
        check(srv.wait(rid, 600000))
    wall = time.perf_counter() - t0
    task_lat, req_lat, flops, ntask = [], [], 0.0, 0
    first, last = None, None
    per_prog = {}
    for k, rid, outs in subs:
        inf = srv.info(rid)
        r, i, d = srv.task_times(rid, inf.ntasks)
        task_lat.append(d - r)
        req_lat.append(inf.t_done - inf.t_submit)
        flops += inf.flops
        ntask += inf.ntasks
        first = inf.t_submit if first is None else min(first, inf.t_submit)
        last = inf.t_done if last is None else max(last, inf.t_done)
        per_prog.setdefault(progs[k].label, []).append(inf.t_done - inf.t_submit)
        srv.release(rid)
    span = (last - first) * 1e-6
    tl = np.concatenate(task_lat) * 1e-3
    rl = np.array(req_lat) * 1e-3
    res = {
        "requests": nreq, "rate_per_s": rate, "statements": ntask,
        "seconds": span, "wall_seconds": wall,
        "tflops": flops / span / 1e12, "requests_per_s": nreq / span,
        "statements_per_s": ntask / span,
        "task_latency_ms": {"p50": float(np.percentile(tl, 50)), "p99": float(np.percentile(tl, 99))},
        "request_latency_ms": {"p50": float(np.percentile(rl, 50)), "p99": float(np.percentile(rl, 99))},
        "per_program_request_ms_p50": {k: float(np.median(v)) * 1e-3 for k, v in sorted(per_prog.items())},
        "usage": srv.usage(),
    }
    return res


def service_times(srv, progs):
    """One request of each program alone: seconds and dense TFLOP/s."""
    out = {}
    for p in progs:
        torch.cuda.synchronize()
        rid, _ = p.submit()
        check(srv.wait(rid, 600000))
        inf = srv.info(rid)
        s = (inf.t_done - inf.t_admit) * 1e-6
        out[p.label] = {"seconds": s, "tflops": inf.flops / s / 1e12, "statements": inf.ntasks,
                        "block_gb": p.block_bytes / 1e9}
        srv.release(rid)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--requests", type=int, default=36)
    ap.add_argument("--rate", type=float, default=2.5)
    ap.add_argument("--mixed", action="store_true", help="add INT8x5 programs to the mix")
    ap.add_argument("--int8x4l6", action="store_true",
                    help="with --mixed: four digits, levels up to 6, instead of INT8x5")
    ap.add_argument("--resident-integrals", action="store_true",
                    help="mixed programs split their integral inputs once, at finalize "
                         "(the amplitudes are split by every request)")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    check(lib().dtcf_init(0))
    precs = (((0, 0), (mixed.I8, 4, 6) if args.int8x4l6 else (mixed.I8, 5)) if args.mixed
             else ((0, 0),))
    log = lambda s: print(s, flush=True)
    srv, progs = build(args.budget_gb, args.slots, precs=precs, log=log,
                       resident_integrals=args.resident_integrals)
    alone = service_times(srv, progs)
    for k, v in alone.items():
        log(f"alone {k}: {v['seconds'] * 1e3:.1f} ms, {v['tflops']:.1f} TF/s, block {v['block_gb']:.2f} GB")
    res = run_stream(srv, progs, args.requests, args.rate, log=log)
    res["alone"] = alone
    log(json.dumps(res))
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)
    for p in progs:
        p.destroy()
    srv.close()


if __name__ == "__main__":
    sys.exit(main())
