"""Block-sparse (D2h) terms as one grouped launch vs the block-by-block
oracle, and the oracle vs the dense contraction of zero-filled tensors."""
import itertools

import numpy as np
import pytest
import torch

import oracle
from dtcf_amd import blocksparse as bsp
from dtcf_amd import mixed

pytestmark = pytest.mark.gpu
TOL = 1e-12


def relerr(got, ref):
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


def meta(t):
    return {g: (t.offset[g], t.shape[g]) for g in t.blocks}


def dense_of(t, buf):
    sizes = [sum(t.space.counts[k]) for k in t.kinds]
    starts = {k: np.concatenate([[0], np.cumsum(t.space.counts[k])]) for k in "ov"}
    out = np.zeros(sizes)
    for g in t.blocks:
        sl = tuple(slice(starts[k][gi], starts[k][gi] + t.space.counts[k][gi])
                   for k, gi in zip(t.kinds, g))
        out[sl] = t.block(g, buf)
    return out


TERMS = [
    # (C kinds, C idx, A kinds, A idx, B kinds, B idx)
    ("oovv", "ijab", "oovv", "ikac", "oovv", "kjcb"),    # ring (config 2 shape)
    ("vvoo", "abij", "vvvv", "abcd", "vvoo", "cdij"),    # ladder (config 1 shape)
    ("vvoo", "abij", "vv", "ac", "vvoo", "cbij"),        # Fock-like (2-D x 4-D)
    ("oo", "ij", "oovv", "ikab", "oovv", "jkab"),        # T1/intermediate-like
    # fastest index external: A (then A and B) gathered transposed in cfg 4
    ("oovv", "ijab", "ovov", "kcjb", "oovv", "ikac"),
    ("oovv", "ijab", "ovov", "kcjb", "ovov", "kcia"),
]


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("term", TERMS)
def test_grouped_small_vs_oracle(dev, term, cfg):
    ck, ci, ak, ai, bk, bi = term
    sp = bsp.Space.seeded(o=16, v=40, seed=3)
    A = bsp.BSTensor(sp, ak, dev).fill_normal(1)
    B = bsp.BSTensor(sp, bk, dev).fill_normal(2)
    Ct = bsp.BSTensor(sp, ck, dev).fill_normal(3)
    Ah, Bh, C0 = A.data.cpu().numpy(), B.data.cpu().numpy(), Ct.data.cpu().numpy()
    t = bsp.BlockTerm(Ct, ci, A, ai, B, bi, cfg=cfg)
    t.run(0.5)
    got = Ct.data.cpu().numpy()
    ref = oracle.bs_contract(meta(Ct), C0.copy(), ci, meta(A), Ah, ai, meta(B), Bh, bi, 0.5)
    assert relerr(got - C0, ref - C0) <= TOL
    # the oracle itself: dense einsum of zero-filled tensors, restricted to C's blocks
    dC = dense_of(Ct, ref) - dense_of(Ct, C0)
    full = 0.5 * np.einsum(f"{ai},{bi}->{ci}", dense_of(A, Ah), dense_of(B, Bh))
    assert relerr(dC, full) <= 1e-12


def test_grouped_deterministic_and_alone_vs_mixed(dev):
    sp = bsp.Space.seeded(o=16, v=40, seed=4)
    A = bsp.BSTensor(sp, "oovv", dev).fill_normal(1)
    B = bsp.BSTensor(sp, "oovv", dev).fill_normal(2)
    outs = []
    for _ in range(3):
        Ct = bsp.BSTensor(sp, "oovv", dev)
        bsp.BlockTerm(Ct, "ijab", A, "ikac", B, "kjcb").run(1.0)
        outs.append(Ct.data.cpu())
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])


@pytest.mark.parametrize("cfg", [None, 4])
def test_grouped_config3_full_size(dev, cfg):
    # config 3: o=64, v=256 over 8 irreps, ~4000 uneven block pairs per term
    sp = bsp.Space.seeded(o=64, v=256, seed=5)
    A = bsp.BSTensor(sp, "oovv", dev).fill_normal(1)
    B = bsp.BSTensor(sp, "oovv", dev).fill_normal(2)
    Ct = bsp.BSTensor(sp, "oovv", dev)
    t = bsp.BlockTerm(Ct, "ijab", A, "ikac", B, "kjcb", cfg=cfg)
    print("pairs", t.npairs, "work", t.nwork, "flops", t.flops, flush=True)
    assert 3000 <= t.npairs <= 5000
    t.run(1.0)
    got = Ct.data.cpu().numpy()
    ref = oracle.bs_contract(meta(Ct), np.zeros(Ct.size), "ijab", meta(A), A.data.cpu().numpy(),
                             "ikac", meta(B), B.data.cpu().numpy(), "kjcb", 1.0)
    assert relerr(got, ref) <= TOL


@pytest.mark.parametrize("gcfg", [0, 3])
@pytest.mark.parametrize("term", TERMS[:2])
def test_irrep_dense_tiles_and_resident_a(dev, term, gcfg):
    """Dense-by-irrep path with each grouped-GEMM tile forced, then a second
    run reusing the gathered A: its increment equals the first bitwise."""
    ck, ci, ak, ai, bk, bi = term
    sp = bsp.Space.seeded(o=16, v=40, seed=6)
    A = bsp.BSTensor(sp, ak, dev).fill_normal(1)
    B = bsp.BSTensor(sp, bk, dev).fill_normal(2)
    Ct = bsp.BSTensor(sp, ck, dev).fill_normal(3)
    Ah, Bh, C0 = A.data.cpu().numpy(), B.data.cpu().numpy(), Ct.data.cpu().numpy()
    t = bsp.BlockTerm(Ct, ci, A, ai, B, bi, cfg=4)
    t._dense.gemm_cfg = gcfg
    t.run(0.5)
    C1 = Ct.data.cpu().numpy()
    ref = oracle.bs_contract(meta(Ct), C0.copy(), ci, meta(A), Ah, ai, meta(B), Bh, bi, 0.5)
    assert relerr(C1 - C0, ref - C0) <= TOL
    Ct.data.copy_(torch.from_numpy(C0))
    t.run(0.5, a_resident=True)
    assert np.array_equal(Ct.data.cpu().numpy(), C1)


# (precision, digits, level) -> bound on the error of C relative to its largest element
MIXED = [((mixed.I8, 5, 6), 5e-10), ((mixed.I8, 4, 5), 1e-7), ((mixed.I8, 3, 4), 2e-5),
         ((mixed.BF16X3, 3, 0), 5e-6), ((mixed.F16X3, 3, 0), 5e-6), ((mixed.F32, 1, 0), 5e-6)]


@pytest.mark.parametrize("mode,bound", MIXED, ids=lambda x: str(x))
@pytest.mark.parametrize("term", TERMS[:2] + TERMS[5:], ids=["ring", "ladder", "both_kxn"])
def test_mixed_dense_by_irrep_vs_fp64(dev, term, mode, bound):
    """The dense-by-irrep form in a mixed mode against its FP64 form on the
    same block-sparse operands; a second run with A resident (slices kept)
    gives the same C bit for bit, and switching back to FP64 gives the FP64
    C bit for bit."""
    ck, ci, ak, ai, bk, bi = term
    sp = bsp.Space.seeded(o=16, v=40, seed=3)
    A = bsp.BSTensor(sp, ak, dev).fill_normal(1)
    B = bsp.BSTensor(sp, bk, dev).fill_normal(2)
    C64 = bsp.BSTensor(sp, ck, dev)
    t = bsp.BlockTerm(C64, ci, A, ai, B, bi, cfg=4)
    t.run(1.0)
    Cm = bsp.BSTensor(sp, ck, dev)
    t.set_precision(*mode)
    t.run(1.0, Cb=Cm.data)
    ref, got = C64.data.cpu().numpy(), Cm.data.cpu().numpy()
    err = relerr(got, ref)
    print(term[1], mode, "rel err", err, flush=True)
    assert err <= bound
    again = torch.zeros_like(Cm.data)
    t.run(1.0, Cb=again, a_resident=True)
    assert torch.equal(again, Cm.data)
    t.set_precision(0)
    back = torch.zeros_like(Cm.data)
    t.run(1.0, Cb=back, a_resident=True)
    assert torch.equal(back, C64.data)


def test_mixed_requires_dense_by_irrep(dev):
    sp = bsp.Space.seeded(o=16, v=40, seed=3)
    A = bsp.BSTensor(sp, "oovv", dev).fill_normal(1)
    t = bsp.BlockTerm(bsp.BSTensor(sp, "oovv", dev), "ijab", A, "ikac", A, "kjcb", cfg=2)
    with pytest.raises(ValueError):
        t.set_precision(mixed.I8, 4, 5)
