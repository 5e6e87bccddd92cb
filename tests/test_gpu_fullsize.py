"""GPU checks at the workloads' full sizes, where the CPU oracle cannot run
whole: the C2 ring step (o=48, v=192, 9216^3 FP64 GEMM, the benchmark's
value) through both step modes and against CPU FP64 dot products of sampled
rows; the GEMM tail split over K (the branch the C2 step takes); the C3
block-sparse ladder (o=64, v=256) against the block-pair oracle on sampled C
blocks; and the UHF T1 / T2 programs at o=64, v=256 against the term-list
oracle evaluated on sampled boxes of the residuals."""
import sys

import numpy as np
import pytest
import torch

import oracle
from dtcf_amd import ops, rrr

pytestmark = pytest.mark.gpu
GEMM_TOL = 1e-12  # max|C - C_ref| / max|C_ref|, FP64


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(got, ref):
    ref = np.asarray(ref)
    return float(np.max(np.abs(np.asarray(got) - ref)) / max(np.max(np.abs(ref)), 1e-300))


def _cus():
    return torch.cuda.get_device_properties(0).multi_processor_count


def _tail_split_shape():
    """m, n of a 128 x 128-tiled GEMM whose tile count is one remainder of
    slots/8 past a whole round of slots (two workgroups per CU): the tail
    tiles are split over K (C2 at 9216^3: 5184 tiles on 512 slots)."""
    slots = 2 * _cus()
    T = slots + max(1, slots // 8)
    tn = max(d for d in range(1, int(T ** 0.5) + 1) if T % d == 0)
    tm = T // tn
    return tm * 128 - 5, tn * 128 - 3, T, slots


@pytest.mark.parametrize("kind", ["dense", "view"])
def test_gemm_tail_split_branch(dev, kind):
    m, n, T, slots = _tail_split_shape()
    k = 1000  # 63 K steps of 16 >= 32: the tail tiles take K slices
    assert T > slots and 0 < T % slots <= slots // 2
    say(f"tail split {kind}: m={m} n={n} k={k} tiles={T} slots={slots}")
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, k))
    B = rng.standard_normal((n, k))
    C0 = rng.standard_normal((m, n))
    ops.set_gemm_config(0)
    try:
        Ad, Bd = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
        if kind == "dense":
            Cd = torch.from_numpy(C0.copy()).to(dev)
            ops.gemm(0, 1, m, n, k, 0.5, Ad, k, Bd, k, Cd, n)
            got = Cd.cpu().numpy()
        else:
            # A and B as views, C a column-major view (the C2 step's CVIEW form)
            av = ops.View(m, k, np.arange(m) * k, np.arange(k))
            bv = ops.View(n, k, np.arange(n) * k, np.arange(k))
            cv = ops.View(m, n, np.arange(m), np.arange(n) * m)
            Cd = torch.from_numpy(np.ascontiguousarray(C0.T).ravel()).to(dev)
            ops.gemm_view(m, n, k, 0.5, Ad, av, Bd, bv, Cd, cv)
            got = Cd.cpu().numpy().reshape(n, m).T
        torch.cuda.synchronize()
    finally:
        ops.set_gemm_config(-1)
    ref = C0 + 0.5 * (A @ B.T)
    err = relerr(got, ref)
    say(f"  relative error {err:.2e}")
    assert err <= GEMM_TOL


def test_c2_ring_full_size_fused_vs_packed_and_sampled_rows(dev):
    """C[i,j,a,b] += sum_kc A[i,k,a,c] B[k,j,c,b] at o=48, v=192 (RRR tiles
    12x12x48x48, vgrid 4x4x4x4): FUSED (the measured mode) against PACKED
    (dense big matrices) everywhere, and both against CPU FP64 sums on 12
    sampled (i, a) rows of C."""
    o, v = 48, 192
    vg = [4, 4, 4, 4]
    rng = np.random.default_rng(48192)
    shape = (o, o, v, v)
    A = rng.standard_normal(shape) * 0.05
    B = rng.standard_normal(shape) * 0.05
    C0 = rng.standard_normal(shape) * 0.05
    addrs = rrr.tile_addresses("cccc", vg)
    At, Bt, Ct = (rrr.tiles_from_dense(x, vg, addrs) for x in (A, B, C0))
    res = {}
    for mode, name in ((rrr.PACKED, "packed"), (rrr.FUSED, "fused")):
        say(f"C2 full size, mode {name}")
        tens = []
        for host in (At, Bt, Ct):
            t = rrr.TileTensor("cccc", list(shape), vg, dev, addrs=addrs)
            t.tiles.copy_(torch.from_numpy(host))
            tens.append(t)
        st = rrr.LocalStep(tens[0], "i,k,a,c", tens[1], "k,j,c,b", tens[2], "i,j,a,b", mode=mode)
        assert st.big_mnk == (o * v, o * v, o * v)
        st.run(1.0)
        torch.cuda.synchronize()
        res[name] = rrr.dense_from_tiles(tens[2].tiles.cpu().numpy(), shape, vg, addrs)
        del st, tens
        torch.cuda.empty_cache()
    d = relerr(res["fused"], res["packed"])
    say(f"  fused vs packed: {d:.2e}")
    assert d <= GEMM_TOL
    picks = [(int(rng.integers(o)), int(rng.integers(v))) for _ in range(12)]
    for i, a in picks:
        ref = C0[i, :, a, :] + np.einsum("kc,kjcb->jb", A[i, :, a, :], B, optimize=True)
        for name in ("fused", "packed"):
            e = relerr(res[name][i, :, a, :], ref)
            assert e <= GEMM_TOL, (name, i, a, e)


def test_c3_ladder_full_size_sampled_blocks(dev):
    """R[a,b,i,j] += 1/2 sum_cd V[a,b,c,d] T[c,d,i,j] over the D2h blocks of
    o=64, v=256 (about 4000 block pairs), as the autotuned path runs it;
    16 C blocks checked against the block-pair oracle."""
    from dtcf_amd import blocksparse as bsp

    sp = bsp.Space.seeded(o=64, v=256, seed=5)
    V = bsp.BSTensor(sp, "vvvv", dev).fill_normal(1)
    T = bsp.BSTensor(sp, "vvoo", dev).fill_normal(2)
    R = bsp.BSTensor(sp, "vvoo", dev).fill_normal(3)
    R0 = R.data.cpu().numpy()
    term = bsp.BlockTerm(R, "abij", V, "abcd", T, "cdij")
    say(f"C3 ladder full size: pairs {term.npairs}, flops {term.flops:.3e}")
    assert 3000 <= term.npairs <= 5000
    term.run(0.5)
    torch.cuda.synchronize()
    got = R.data.cpu().numpy()
    Vh, Th = V.data.cpu().numpy(), T.data.cpu().numpy()
    meta = lambda t: {g: (t.offset[g], t.shape[g]) for g in t.blocks}
    rng = np.random.default_rng(9)
    blocks = list(R.blocks)
    pick = [blocks[i] for i in rng.choice(len(blocks), 16, replace=False)]
    cm = {g: (R.offset[g], R.shape[g]) for g in pick}
    ref = oracle.bs_contract(cm, R0.copy(), "abij", meta(V), Vh, "abcd", meta(T), Th, "cdij", 0.5)
    for g in pick:
        off, shp = R.offset[g], R.shape[g]
        n = int(np.prod(shp))
        e = relerr(got[off:off + n] - R0[off:off + n], ref[off:off + n] - R0[off:off + n])
        assert e <= GEMM_TOL, (g, e)


def _statements(tl):
    out = []
    for s in tl.statements:
        if s.kind == "mul":
            out.append((s.lhs.name, "mul", s.lhs.idx, [(t.name, t.idx) for t in s.terms]))
        else:
            out.append((s.lhs.name, "sum", s.lhs.idx, [(t.name, t.idx, t.coef) for t in s.terms]))
    return out


@pytest.mark.parametrize("which,prec", [("t1", 0), ("t2", 0), ("t2", "int8x5")])
def test_uhf_program_o64_v256_sampled(dev, which, prec):
    """A whole T1 or T2 residual program at o=64, v=256 (alpha and beta halves
    Oa = Ob = 32, Va = Vb = 128) on the server, its residuals compared on
    sampled boxes (two values of each virtual index, every occupied one)
    with the term-list oracle evaluated on those boxes only.  FP64 to 1e-12;
    the INT8x5 program (contractions with m n / (m + n) >= 256 on digit
    slices) to 1e-9."""
    from dtcf_amd import mixed, termlist
    from dtcf_amd import serve as S

    tl = termlist.load(which)
    print(f"UHF {which} program o=64 v=256 prec={prec}", flush=True)
    with S.Server(48 << 30, nslots=8) as srv:
        sysm = S.System(srv, 64, 256, [tl], seed=64256)
        kw = {} if prec == 0 else dict(prec=mixed.I8, nslice=5)
        prog = S.Program(srv, tl, sysm, **kw)
        got = prog.run()
        prog.destroy()
        shapes = {n: tuple(sysm.sizes[r] for r in a.ranges) for n, a in tl.arrays.items()}
        rng = np.random.default_rng(7)
        want = {}
        for n in tl.outputs():
            shp = shapes[n]
            want[n] = [sorted(rng.choice(shp[0], 2, replace=False).tolist())]
            if len(shp) == 4:
                want[n] += [sorted(rng.choice(shp[1], 2, replace=False).tolist()), None, None]
            else:
                want[n] += [None]
        cache = {}

        def get(name):
            if name not in cache:
                cache[name] = sysm.host(name)
            return cache[name]

        ref = oracle.termlist_run_sampled(_statements(tl), get, shapes, want)
    tol = 1e-12 if prec == 0 else 1e-9
    for n, box in want.items():
        idx = [np.arange(shapes[n][d]) if b is None else np.asarray(b) for d, b in enumerate(box)]
        g = got[n][np.ix_(*idx)]
        e = relerr(g, ref[n])
        say(f"  {n}: box {ref[n].shape}, relative error {e:.2e}")
        assert e <= tol, (n, e)
