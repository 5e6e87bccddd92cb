#!/usr/bin/env python3
"""Benchmark: one rank's local contraction step of the ring term on MI355X.

Workload (single-rank SUMMA-step share, FP64):
    C[i,j,a,b] += sum_{k,c} A[i,k,a,c] * B[k,j,c,b],   o = 48, v = 192
A, B, C are held as RRR tiles (vgrid 4x4x4x4 -> 256 tiles of 12x12x48x48
each per tensor, generate_data order), seeded N(0, 0.05^2).  A "step" is
one dtcf_step_run: the transposes into GEMM form, the GEMM and the C
update, with every input already resident in HBM.  The metric is FP64
contraction TFLOP/s = 2*(o*v)^3 / step time.

Also measured (extra keys of the same line): full UHF CCSD iterations, FP64
against mixed-precision residual programs (time per iteration, energy error
per component), the precision study (FP64 vs FP32 / BF16 / FP16 / INT8 digit
splits, ladder + ring at o=64, v=256, errors per term), the block-sparse D2h
terms as grouped launches, and the served UHF CCSD stream (throughput and
p50 / p99 latency).  --no-extra skips them.

Prints ONE JSON line.  --gpus N > 1 exits: one rank's share runs on one GPU.
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "contraction TFLOP/s FP64 vs mixed at stated error; task p50/p99 latency"
FP64_PEAK_TF = 78.6      # MI355X dense FP64 MFMA
HBM_PEAK_GBS = 8000.0
# HBM bytes per GEMM launch come from rocprofv3 --pmc runs of this script
# (FETCH_SIZE, then WRITE_SIZE, scaled by a plain copy's counts), recorded in
# profiles/traffic.json with the kernel's average duration in those runs.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def sig(x, n=4):
    """Round to n significant digits (keeps tiny errors visible in the JSON)."""
    return float(f"{x:.{n}g}") if isinstance(x, float) else x


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(o, v, vg):
    from dtcf_amd import rrr, synth

    shape = (o, o, v, v)
    n = int(np.prod(shape))
    addrs = rrr.tile_addresses("cccc", vg)
    out = []
    for seed in (1, 2, 3):
        dense = synth.normal(n, seed, 0.05).reshape(shape)
        out.append(rrr.tiles_from_dense(dense, vg, addrs))
        del dense
    return out, addrs


def host_cores():
    """Cores this process may use, counted at run time: the physical cores
    (distinct (package, core) pairs of /proc/cpuinfo) among the CPUs of its
    affinity mask, capped by the cgroup CPU quota when one is set.  Returns
    (count, description)."""
    allowed = sorted(os.sched_getaffinity(0))
    phys, cur = {}, {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if "processor" in cur:
                        phys[cur["processor"]] = (cur.get("physical id", "0"), cur.get("core id", cur["processor"]))
                    cur = {}
                    continue
                k, val = (x.strip() for x in line.split(":", 1))
                cur[k] = val
        if "processor" in cur:
            phys[cur["processor"]] = (cur.get("physical id", "0"), cur.get("core id", cur["processor"]))
    except OSError:
        pass
    cores = {phys.get(str(c), ("0", str(c))) for c in allowed}
    n = max(1, len(cores))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()
            if q != "max":
                quota = max(1, int(int(q) / int(per)))
    except (OSError, ValueError):
        pass
    desc = (f"{len(allowed)} logical CPUs in the affinity mask, {len(cores)} physical cores"
            + (f", cgroup quota {quota} CPUs" if quota else ""))
    return (min(n, quota) if quota else n), desc


def cpu_baseline(A_t, B_t, C_t, o, v, vg, target_s=12.0):
    """The restated reference CPU path (oracle, OpenMP) on this host, the
    reference's tr_time / comp_time split (contraction.cc:2923-2926) kept:
    tile transposes (helper.cc:275-522) and create_big_matrix packs
    (transpose_helper.cc:20-171) of A, B and C in full, the big dgemm on a
    row slab of bigA sized to ~target_s and extrapolated to the full M, then
    revert_big_matrix of C (transpose_helper.cc:331-409) and the inverse
    transposes of its tiles (contraction.cc:2817-2830) in full."""
    import oracle

    threads, core_desc = host_cores()
    br = [o // vg[0], o // vg[1], v // vg[2], v // vg[3]]
    bs = int(np.prod(br))
    ntiles = len(A_t) // bs
    M = N = K = o * v
    pm = oracle.permutation_maps(list("ikac"), list("kjcb"), list("ijab"),
                                 dict(i=br[0], k=br[1], a=br[2], c=br[3], j=br[1], b=br[3]))
    addrs = np.array([np.unravel_index(t, vg) for t in range(ntiles)], dtype=np.int32)
    t0 = time.perf_counter()
    pa = oracle.rrr_transpose_tiles(A_t, pm["p_map_A"], br, threads)
    pb = oracle.rrr_transpose_tiles(B_t, pm["p_map_B"], br, threads)
    pc = oracle.rrr_transpose_tiles(C_t, pm["p_map_C"], br, threads)
    t_tr = time.perf_counter() - t0

    def order_for(pmap, nrow):
        perm = np.zeros_like(addrs)
        for d in range(4):
            perm[:, pmap[d]] = addrs[:, d]
        return oracle.big_matrix_order(vg, pmap, perm, nrow)

    t0 = time.perf_counter()
    oa, nra, nca = order_for(pm["p_map_A"], 2)
    ob, nrb, ncb = order_for(pm["p_map_B"], 2)
    oc, nrc, ncc = order_for(pm["p_map_C"], 2)
    rpb = br[0] * br[2]
    bigA = oracle.big_matrix_pack(pa, oa, nra, nca, rpb, rpb)
    bigB = oracle.big_matrix_pack(pb, ob, nrb, ncb, rpb, rpb)
    bigC = oracle.big_matrix_pack(pc, oc, nrc, ncc, rpb, rpb)
    t_pack = time.perf_counter() - t0
    del pa, pb, pc
    # calibrate on a thin slab, then run a slab sized to target_s
    rows = 64
    cs = np.zeros(rows * N)
    t0 = time.perf_counter()
    oracle.dgemm_nt_blocked(rows, N, K, 1.0, bigA[:rows].ravel(), K, bigB.ravel(), K, cs, N,
                            threads)
    rate = 2.0 * rows * N * K / (time.perf_counter() - t0)
    rows = int(min(M, max(64, target_s * rate / (2.0 * N * K))))
    rows -= rows % 64
    rows = max(rows, 64)
    cs = bigC[:rows].ravel().copy()
    t0 = time.perf_counter()
    oracle.dgemm_nt_blocked(rows, N, K, 1.0, np.ascontiguousarray(bigA[:rows]).ravel(), K,
                            bigB.ravel(), K, cs, N, threads)
    t_gemm_slab = time.perf_counter() - t0
    t_gemm = t_gemm_slab * M / rows
    del bigA, bigB
    # revert_big_matrix of C, then the inverse transposes of its tiles
    t0 = time.perf_counter()
    tiles_c = oracle.big_matrix_unpack(bigC, oc, nrc, ncc, rpb, rpb, ntiles)
    t_unpack = time.perf_counter() - t0
    del bigC
    inv = [0] * 4
    for d in range(4):
        inv[pm["p_map_C"][d]] = d
    br_perm = [0] * 4
    for d in range(4):
        br_perm[pm["p_map_C"][d]] = br[d]
    t0 = time.perf_counter()
    oracle.rrr_transpose_tiles(tiles_c, inv, br_perm, threads)
    t_untr = time.perf_counter() - t0
    t_total = t_tr + t_pack + t_gemm + t_unpack + t_untr
    flops = 2.0 * M * N * K
    return {
        "value": round(flops / t_total / 1e12, 5),
        "unit": "TFLOP/s",
        "cores": threads,
        "kind": "port",
        "host": core_desc,
        "seconds": {"transpose": round(t_tr, 3), "pack": round(t_pack, 3),
                    "dgemm_extrapolated": round(t_gemm, 3), "unpack_c": round(t_unpack, 3),
                    "inverse_transpose_c": round(t_untr, 3)},
        "sample": (f"full tile transposes and big-matrix packs of A, B, C; blocked OpenMP dgemm "
                   f"on a {rows}x{N}x{K} row slab of the {M}x{N}x{K} GEMM ({t_gemm_slab:.2f}s) "
                   f"extrapolated to all rows; full unpack and inverse transposes of C; "
                   f"{threads} threads"),
    }


def traffic_record(mode, o, v, live_ms):
    """HBM bytes per launch of the step's GEMM from profiles/traffic.json
    (FETCH_SIZE x fetch_scale + WRITE_SIZE x write_scale, KiB per dispatch)
    for this workload, with the profiled kernel's average duration beside
    the live one: a kernel that has changed since its profile shows as a
    duration mismatch (traffic_stale)."""
    try:
        with open(TRAFFIC_FILE) as f:
            recs = json.load(f)
    except (OSError, ValueError):
        return {}
    for key, r in recs.items():
        w = r.get("workload", {})
        if (w.get("o"), w.get("v"), w.get("mode")) != (o, v, mode):
            continue
        traffic = 1024.0 * (r["fetch_kib_per_dispatch"] * r["fetch_scale"] +
                            r["write_kib_per_dispatch"] * r["write_scale"])
        prof_ms = r["avg_duration_us"] / 1e3
        return {"traffic": round(traffic, -6), "traffic_source": f"profiles/traffic.json:{key}",
                "traffic_profile_kernel_ms": round(prof_ms, 4),
                "traffic_stale": abs(prof_ms - live_ms) > 0.05 * live_ms}
    return {}


def leg_ccsd(log, o=96, v=384, iters=3):
    """Full UHF CCSD iterations (T1 + T2 residual programs, FP64 amplitude
    update, singles-inclusive correlation energy): FP64 against mixed
    residual programs from the same amplitudes.  The system is generated as
    the precision study's (eps_occ ~ U[-2,-0.5], eps_vir ~ U[0.2,3],
    antisymmetrised normal integrals) with sigma = 0.5 / v, a scale at
    which the Jacobi iteration converges; with sigma = 0.05 it diverges, and
    that case is reported for the first iteration from the first-order
    amplitudes at o=64, v=256."""
    from dtcf_amd import ccsd
    res = ccsd.study(o, v, iters, modes=("fp64", "int8x4/L5", "int8x4/L4", "int8x3/L4"),
                     sigma=0.5 / v, log=lambda s_: log("[ccsd] " + s_))
    out = {"workload": f"UHF CCSD iterations o={o} v={v} (Oa=Ob={o // 2}, Va=Vb={v // 2}), "
                       f"sigma={0.5 / v:.3g}, {iters} iterations from the first-order amplitudes, "
                       f"every mode from the FP64 iteration's amplitudes, one server slot",
           "flops_per_iteration": res["flops_per_iteration"], "modes": {}}
    for m, r in res["modes"].items():
        out["modes"][m] = {"seconds_per_iteration": round(r["seconds_per_iteration"], 4),
                           "tflops_residuals": round(r["tflops_residuals"], 2),
                           "speedup": round(r["speedup"], 3),
                           "energy": [round(x, 10) for x in r["energy"]],
                           "max_energy_err": {k: sig(x) for k, x in r["max_energy_err"].items()},
                           "max_residual_rel_err": {k: sig(x) for k, x in
                                                    r["max_residual_rel_err"].items()}}
    ok = [(m, r) for m, r in res["modes"].items()
          if m != "fp64" and r["max_energy_err"]["E"] <= 1e-7]
    if ok:
        m, r = max(ok, key=lambda x: x[1]["speedup"])
        out["best_within_1e-7_Eh"] = {"mode": m, "speedup": round(r["speedup"], 3),
                                      "max_energy_err": sig(r["max_energy_err"]["E"])}
    import torch
    torch.cuda.empty_cache()
    first = ccsd.study(64, 256, 1, modes=("fp64", "int8x5", "int8x4/L6", "int8x3/L4"), sigma=0.05,
                       log=lambda s_: log("[ccsd sigma=0.05] " + s_))
    out["sigma_0.05_first_iteration_o64v256"] = {
        m: {"speedup": round(r["speedup"], 3), "energy": r["energy"][0],
            "energy_err": sig(r["max_energy_err"]["E"])} for m, r in first["modes"].items()}
    torch.cuda.empty_cache()
    return out


def leg_thin(log):
    """The thin FP64 products of the UHF iteration (one side 48 or 192 wide,
    short K) on the streaming thin kernel and on 64 x 64 tiles."""
    from dtcf_amd import thinbench

    log("[bench] thin FP64 products (streaming kernel vs 64x64 tiles)")
    return thinbench.run(reps=5)


def leg_dropin(log):
    """The drop-in entry points at the reference's call shapes, from C++
    through the exported symbols (dtcf_amd/dropin_bench, built with the
    libraries): local_contraction at 9216^3 with host operands, and the o=8,
    v=32 ladder's 1024 tile tasks issued as the reference's task_mul does
    (host tiles) and as TaskFunc tasks over device tensors."""
    import subprocess
    exe = os.path.join(ROOT, "dtcf_amd", "dropin_bench")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError(f"dropin_bench failed ({r.returncode}): {r.stdout[-300:]} {r.stderr[-300:]}")
    out = json.loads(r.stdout.strip().splitlines()[-1])
    log(f"[dropin] {json.dumps(out)}")
    return out


def leg_precision(log):
    """Precision study (ladder + ring, o=64, v=256): FP64 against the mixed
    modes that bracket the 1e-7 Eh target."""
    from dtcf_amd import mixed, precision
    variants = [v for v in precision.VARIANTS
                if v[0] in ("fp64", "fp32", "bf16x3", "int8x4", "int8x4/L6", "int8x4/L6, ring L5",
                            "int8x5")]
    res = precision.study(64, 256, variants=variants, reps=2, log=lambda s_: log("[precision] " + s_))
    keep = ("name", "seconds", "seconds_integrals_resident", "tflops", "energy_err", "max_rel_err",
            "speedup", "speedup_integrals_resident", "products")

    def terms(x):
        return {t: {k: sig(val) for k, val in d.items()} for t, d in x.get("terms", {}).items()}

    out = {"workload": "ladder + ring, o=64 v=256, one CCD-like iteration", "variants": [
        dict({k: sig(x[k]) for k in keep}, terms=terms(x)) for x in res["variants"]]}
    ok = [x for x in res["variants"] if x["name"] != "fp64" and x["energy_err"] <= 1e-7]
    if ok:
        b = max(ok, key=lambda x: x["speedup_integrals_resident"])
        out["best_within_1e-7_Eh"] = {"name": b["name"], "speedup": round(b["speedup"], 3),
                                      "speedup_integrals_resident": round(b["speedup_integrals_resident"], 3),
                                      "energy_err": sig(b["energy_err"])}
    import torch
    torch.cuda.empty_cache()
    return out


def leg_blocksparse(dev, log, steps=5):
    """Block-sparse D2h terms (o=64, v=256, 8 irreps): one grouped persistent
    launch per term over its symmetry-allowed block pairs; then the
    dense-by-irrep form in FP64 against INT8 slice modes (time, speedup and
    the error of C against FP64, per term)."""
    from dtcf_amd import blocksparse as bsp
    out = bsp.bench_terms(dev, steps=steps)
    for k, v in out.items():
        log(f"[blocksparse] {k}: {v}")
    for v in out.values():
        v["frac_fp64_peak"] = round(v["tflops"] / FP64_PEAK_TF, 4)
    mx = bsp.bench_mixed_terms(dev, steps=steps, modes=("int8x4/L5", "int8x3/L4"))
    for k, v in mx.items():
        log(f"[blocksparse mixed] {k}: {v}")
    out["mixed_dense_by_irrep"] = mx
    return out


def leg_symmetric(dev, log, o=48, v=192, steps=3):
    """The C2 ring step with A symmetric in (i,k) ('aacc', vgrid 4x4x4x4):
    the stored tiles (addr0 >= addr1) and the bounced copies of the others
    (the stored tile's data under the swapped address), so A has two bounce
    classes.  Times LITERAL, PACKED and FUSED, and checks each against the
    FUSED step on the dense tiles of the same tensor."""
    import torch

    from dtcf_amd import mixed, rrr

    vg = [4, 4, 4, 4]
    br = [o // 4, o // 4, v // 4, v // 4]
    n = o * o * v * v
    X = mixed.fill_normal(n, 11, 0.05, dev).view(o, o, v, v)
    full = X + X.transpose(0, 1)
    del X
    # dense tiles [a0][a1][a2][a3][block], row-major addresses
    dense = full.view(4, br[0], 4, br[1], 4, br[2], 4, br[3]).permute(0, 2, 4, 6, 1, 3, 5, 7) \
        .contiguous().view(4, 4, 4, 4, -1)
    del full
    addrs = np.array(list(itertools.product(range(4), repeat=4)), dtype=np.int32)
    src = torch.tensor([[a1, a0, a2, a3] if a0 < a1 else [a0, a1, a2, a3] for a0, a1, a2, a3 in addrs],
                       device=dev, dtype=torch.long)
    sym_tiles = dense[src[:, 0], src[:, 1], src[:, 2], src[:, 3]].reshape(-1)
    tA_sym = rrr.TileTensor("aacc", [o, o, v, v], vg, dev, data=sym_tiles, addrs=addrs)
    tA_dense = rrr.TileTensor("cccc", [o, o, v, v], vg, dev, data=dense.reshape(-1), addrs=addrs)
    tB = rrr.TileTensor("cccc", [o, o, v, v], vg, dev, data=mixed.fill_normal(n, 12, 0.05, dev),
                        addrs=addrs)
    C0 = mixed.fill_normal(n, 13, 0.05, dev)
    tC = rrr.TileTensor("cccc", [o, o, v, v], vg, dev, data=C0.clone(), addrs=addrs)
    ref = rrr.LocalStep(tA_dense, "i,k,a,c", tB, "k,j,c,b", tC, "i,j,a,b", mode=rrr.FUSED)
    ref.run(1.0)
    torch.cuda.synchronize()
    want = tC.tiles.clone()
    del ref
    out = {"workload": f"ring step o={o} v={v} with A symmetric in (i,k) ('aacc'), 160 stored "
                       f"tiles + 96 bounced copies", "modes": {}}
    for name, mode in (("literal", rrr.LITERAL), ("packed", rrr.PACKED), ("fused", rrr.FUSED)):
        st = rrr.LocalStep(tA_sym, "i,k,a,c", tB, "k,j,c,b", tC, "i,j,a,b", mode=mode)
        tC.tiles.copy_(C0)
        st.run(1.0)
        torch.cuda.synchronize()
        err = float((tC.tiles - want).abs().max() / want.abs().max())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st.run(1.0)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        out["modes"][name] = {"ms_per_step": round(dt * 1e3, 3),
                              "tflops": round(st.flops / dt / 1e12, 3),
                              "max_rel_diff_vs_dense_fused": err}
        log(f"[bench] symmetric A, {name}: {dt*1e3:.3f} ms/step, rel diff {err:.2e}")
        del st
        torch.cuda.synchronize()
    del tA_sym, tA_dense, tB, tC, dense, sym_tiles, C0, want
    torch.cuda.empty_cache()
    return out


def leg_c1(dev, log, reps=20):
    """C1 CCD ladder R[a,b,i,j] += 0.5 V[a,b,c,d] T[c,d,i,j] at o=8, v=32
    (ccsd-t2.cc:57-61) through the DLTC term path: latency of the whole
    term as one tile (wbO=8, wbV=32: one GEMM) and of the tile-task form
    (wbO=4, wbV=8: one task per (external tile, summed tile), as DLTC's
    finest granularity issues them), inputs antisymmetrised on the device."""
    import itertools
    import torch
    from dtcf_amd import dltc, mixed
    o, v = 8, 32
    V = mixed.antisymmetrize4(mixed.fill_normal(v ** 4, 1, 0.05, dev), (v, v, v, v))
    T = mixed.antisymmetrize4(mixed.fill_normal(v * v * o * o, 2, 0.05, dev), (v, v, o, o))
    h = {f"h{i}": i - 1 for i in range(1, 13)}
    p = {f"p{i}": 11 + i for i in range(1, 13)}
    c_ids = [p["p3"], p["p4"], h["h1"], h["h2"]]
    a_ids = [p["p5"], p["p6"], h["h1"], h["h2"]]
    b_ids = [p["p3"], p["p4"], p["p5"], p["p6"]]
    flops = 2.0 * v ** 4 * o ** 2
    out = {"workload": "C1 CCD ladder o=8 v=32, FP64", "flops": flops}
    R = torch.zeros(v * v * o * o, dtype=torch.float64, device=dev)
    term = dltc.Term(dltc.make_space(o, v, 8, 32), c_ids, R, a_ids, T, b_ids, V, 0.5)
    term.run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        term.run()
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / reps
    out["one_tile_us"] = round(us, 2)
    out["one_tile_tflops"] = round(flops / us / 1e6, 3)
    wbO, wbV = 4, 8
    term2 = dltc.Term(dltc.make_space(o, v, wbO, wbV), c_ids, R, a_ids, T, b_ids, V, 0.5)
    ext, sumi = c_ids, [p["p5"], p["p6"]]
    spO, spV = o // wbO, v // wbV
    boxes = []
    for e in itertools.product(range(spV), range(spV), range(spO), range(spO)):
        for sm in itertools.product(range(spV), range(spV)):
            lb, ub = [0] * 24, [0] * 24
            for x, val in zip(ext + sumi, e + sm):
                lb[x], ub[x] = val, val + 1
            boxes.append((lb, ub))
    torch.cuda.synchronize()
    t = time.perf_counter()
    for lb, ub in boxes:
        term2.task(lb, ub)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    out["tasks"] = len(boxes)
    out["tasks_ms"] = round(dt * 1e3, 3)
    out["us_per_task"] = round(dt * 1e6 / len(boxes), 2)

    # the same tasks recorded once as a HIP graph (the 64 output tiles'
    # chains of 16 sum-tile tasks over 4 branches) and replayed
    per = len(boxes) // (spV * spV * spO * spO)

    def chain(i):
        return lambda: [term2.task(lb, ub) for lb, ub in boxes[i:i + per]]

    graph = dltc.TaskGraph([chain(i) for i in range(0, len(boxes), per)], branches=4)
    graph.replay()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        graph.replay()
    e1.record()
    torch.cuda.synchronize()
    gms = e0.elapsed_time(e1) / reps
    out["tasks_graph_ms"] = round(gms, 3)
    out["us_per_task_graph"] = round(gms * 1e3 / len(boxes), 2)
    del graph
    log(f"[c1] {json.dumps(out)}")
    del V, T, R
    torch.cuda.empty_cache()
    return out


def leg_serving(log, requests=36, rate=2.5):
    """Served CCSD iteration stream: UHF T1 + T2 programs over o/v in
    {(32,128), (64,256), (96,384)}, Poisson arrivals, 200 GB HBM budget."""
    from dtcf_amd import mixed, servebench
    srv, allp = servebench.build(200.0, 8, precs=((0, 0), (mixed.I8, 5)),
                                 log=lambda s_: log("[serving] " + s_))
    progs = [p for p in allp if p.label.endswith("fp64")]
    try:
        alone = servebench.service_times(srv, allp)
        res = servebench.run_stream(srv, progs, requests, rate)
        sat = servebench.run_stream(srv, progs, 18, 0.0, seed=2)
        # the same programs plus their INT8x5 forms (contractions with
        # m n / (m + n) >= 256 on digit slices, the rest FP64), all at once
        mix = servebench.run_stream(srv, allp, 24, 0.0, seed=3)
        # error of the INT8x5 residuals against the FP64 program's, o=64 v=256
        byl = {p.label: p for p in allp}
        err = {}
        for which in ("t1", "t2"):
            r64 = byl[f"{which}/o64v256/fp64"].run()
            rmx = byl[f"{which}/o64v256/int8x5"].run()
            err[which] = max(float(np.abs(rmx[n] - r64[n]).max() / max(np.abs(r64[n]).max(), 1e-300))
                             for n in r64)
    finally:
        for p in allp:
            p.destroy()
        srv.close()
    out = {"workload": "UHF T1+T2 residual programs (72 + 306 statements), systems o/v in "
                       "{32/128, 64/256, 96/384} (alpha = beta = half), FP64, 8 slots, 200 GB budget",
           "poisson": {k: res[k] for k in ("requests", "rate_per_s", "statements", "seconds", "tflops",
                                          "requests_per_s", "statements_per_s", "task_latency_ms",
                                          "request_latency_ms")},
           "saturated": {k: sat[k] for k in ("requests", "statements", "seconds", "tflops",
                                            "requests_per_s", "task_latency_ms", "request_latency_ms")},
           "saturated_with_int8x5": {k: mix[k] for k in ("requests", "statements", "seconds", "tflops",
                                                        "requests_per_s", "task_latency_ms",
                                                        "request_latency_ms")},
           "int8x5_max_rel_err_o64v256": {k: sig(v) for k, v in err.items()},
           "alone_ms": {k: round(v["seconds"] * 1e3, 3) for k, v in alone.items()},
           "alone_tflops": {k: round(v["tflops"], 2) for k, v in alone.items()},
           "peak_reserved_gb": round(res["usage"]["peak_reserved"] / 1e9, 2),
           "resident_gb": round(res["usage"]["resident"] / 1e9, 2)}
    log(f"[serving] {json.dumps(out)}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", choices=["literal", "packed", "fused", "best"], default="best")
    ap.add_argument("--o", type=int, default=48)
    ap.add_argument("--v", type=int, default=192)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the precision, block-sparse and serving legs")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 2 GiB device copy (for PMC traffic scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: one rank's share runs on one GPU; --gpus {args.gpus} is not supported "
              "(the inter-rank SUMMA and redistribution are outside this library)", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    import torch

    from dtcf_amd import _lib, rrr

    _lib.check(_lib.lib().dtcf_init(0))
    dev = torch.device("cuda:0")
    o, v = args.o, args.v
    vg = [4, 4, 4, 4]
    log(f"[bench] generating inputs o={o} v={v}")
    (A_t, B_t, C_t), addrs = make_inputs(o, v, vg)
    tens = []
    for host in (A_t, B_t, C_t):
        t = rrr.TileTensor("cccc", [o, o, v, v], vg, dev, addrs=addrs)
        t.tiles.copy_(torch.from_numpy(host))
        tens.append(t)
    tA, tB, tC = tens
    torch.cuda.synchronize()

    modes = {"literal": rrr.LITERAL, "packed": rrr.PACKED, "fused": rrr.FUSED}
    per_mode = {}
    K, W = args.steps, args.warmup
    chosen = None
    for name in (["literal", "packed", "fused"] if args.mode == "best" else [args.mode]):
        st = rrr.LocalStep(tA, "i,k,a,c", tB, "k,j,c,b", tC, "i,j,a,b", mode=modes[name])
        for _ in range(W):
            st.run(1.0)
        st.enable_timing(K)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(K):
            st.run(1.0)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / K
        kt = st.kernel_times_ms(K)
        per_mode[name] = dict(ms_per_step=dt * 1e3, tflops=st.flops / dt / 1e12,
                              gemm_ms=float(np.mean(kt)) if kt else None, flops=st.flops,
                              big_mnk=st.big_mnk)
        log(f"[bench] mode {name}: {dt*1e3:.3f} ms/step, {st.flops/dt/1e12:.2f} TFLOP/s, "
            f"GEMM kernel {per_mode[name]['gemm_ms']:.3f} ms")
        del st
        torch.cuda.synchronize()
    chosen = max(per_mode, key=lambda k: per_mode[k]["tflops"])
    best = per_mode[chosen]

    if args.copy_calibration:
        x = torch.empty(1 << 28, dtype=torch.float64, device=dev)
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        log("[bench] copy calibration: 2 GiB read + 2 GiB written by one copy kernel")

    M, N, Kd = best["big_mnk"]
    gemm_s = best["gemm_ms"] / 1e3
    achieved = 2.0 * M * N * Kd / gemm_s / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK_TF,
                "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TF, 4), "traffic": None,
                "kernel": "gemm_f64_kernel (FP64 MFMA 16x16x4)",
                "kernel_ms": round(best["gemm_ms"], 4)}
    roofline.update(traffic_record(chosen, o, v, best["gemm_ms"]))

    del tA, tB, tC, tens
    torch.cuda.empty_cache()
    extra = {}
    if not args.no_extra:
        for key, fn in (("symmetric", lambda: leg_symmetric(dev, log)),
                        ("c1", lambda: leg_c1(dev, log)),
                        ("dropin", lambda: leg_dropin(log)),
                        ("ccsd", lambda: leg_ccsd(log)),
                        ("thin", lambda: leg_thin(log)),
                        ("precision", lambda: leg_precision(log)),
                        ("blocksparse", lambda: leg_blocksparse(dev, log)),
                        ("serving", lambda: leg_serving(log))):
            try:
                extra[key] = fn()
            except Exception as e:  # a failed side leg must not hide the main line
                log(f"[bench] {key} leg failed: {e!r}")
                extra[key] = {"error": repr(e)}

    cpu = None
    if rank == 0 and not args.no_cpu:
        log("[bench] CPU baseline (restated reference path, oracle)")
        cpu = cpu_baseline(A_t, B_t, C_t, o, v, vg)

    out = {
        "metric": METRIC,
        "value": round(best["tflops"], 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(best["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"ring term C[i,j,a,b] += sum_kc A[i,k,a,c] B[k,j,c,b], o={o} "
                               f"v={v}, RRR tiles vgrid 4x4x4x4, single-rank SUMMA-step share",
                   "o": o, "v": v, "mode": chosen,
                   "modes": {k: {"ms_per_step": round(x["ms_per_step"], 4),
                                 "tflops": round(x["tflops"], 4),
                                 "gemm_ms": round(x["gemm_ms"], 4)} for k, x in per_mode.items()}},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    if rank == 0:
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
