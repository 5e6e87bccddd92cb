"""GPU parity of the core kernels against the CPU oracle (bit-exact for
permutes and packs; normwise relative 1e-12 for FP64 GEMM)."""
import itertools
import sys

import numpy as np
import pytest
import torch

import oracle
from dtcf_amd import ops, rrr

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(7)
GEMM_TOL = 1e-12  # max|C - C_ref| / max|C_ref|, FP64


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(got, ref):
    ref = np.asarray(ref)
    return float(np.max(np.abs(np.asarray(got) - ref)) / max(np.max(np.abs(ref)), 1e-300))


@pytest.mark.parametrize("w", [[5, 7], [32, 33], [3, 4, 5], [12, 12, 48, 48], [2, 3, 4, 5, 3],
                               [1, 9, 1, 4], [2, 3, 2, 4, 3, 2]])
def test_permute_tiles_bit_exact(dev, w):
    say("permute", w)
    n = int(np.prod(w))
    ntiles = 3
    x = RNG.standard_normal(n * ntiles)
    xd = torch.from_numpy(x).to(dev)
    perms = list(itertools.permutations(range(len(w))))
    if len(perms) > 12:
        perms = [perms[i] for i in RNG.choice(len(perms), 12, replace=False)]
    for pm in perms:
        out = ops.rrr_transpose(xd, list(pm), w, ntiles=ntiles)
        ref = np.concatenate([oracle.rrr_transpose(x[t * n:(t + 1) * n], pm, w)
                              for t in range(ntiles)])
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), ref), pm


def test_permute_all_dltc_perms_bit_exact(dev):
    shp = [5, 6, 7, 9]
    n = int(np.prod(shp))
    x = RNG.standard_normal(n)
    xd = torch.from_numpy(x).to(dev)
    for name in oracle.DLTC_4D_NAMES:
        s = ["ijkl".index(ch) for ch in name]
        out = ops.permute_tiles(xd, shp, s)
        assert np.array_equal(out.cpu().numpy(), oracle.permute_by_source(x, s, shp)), name


def test_permute_copy_offsets_and_accumulate(dev):
    # task_sum form: out_tile += coef * perm(in_tile) for a list of tiles
    w = [4, 6, 5]
    bs = 120
    nt = 5
    x = RNG.standard_normal(bs * nt)
    y0 = RNG.standard_normal(bs * 8)
    src = np.array([4, 0, 2, 1, 3]) * bs
    dst = np.array([7, 1, 0, 5, 2]) * bs
    s = [2, 0, 1]
    ext = [w[i] for i in s]
    istr_in = [30, 5, 1]
    istr = [istr_in[i] for i in s]
    ostr = [ext[1] * ext[2], ext[2], 1]
    xd = torch.from_numpy(x).to(dev)
    yd = torch.from_numpy(y0.copy()).to(dev)
    ops.permute_copy(xd, yd, ext, istr, ostr, ntiles=nt,
                     src_off=torch.from_numpy(src).to(dev), dst_off=torch.from_numpy(dst).to(dev),
                     alpha=0.5, beta=1.0)
    ref = y0.copy()
    for t in range(nt):
        p = oracle.permute_by_source(x[src[t]:src[t] + bs], s, w)
        ref[dst[t]:dst[t] + bs] = ref[dst[t]:dst[t] + bs] + 0.5 * p
    assert np.array_equal(yd.cpu().numpy(), ref)


def test_mfma_layout_single_block(dev):
    # one 16x16x4 product: exposes any mistake in the MFMA lane layout
    A = np.arange(64, dtype=np.float64).reshape(16, 4) / 7.0
    B = (np.arange(64, dtype=np.float64).reshape(4, 16) - 30.0) / 3.0
    Cm = np.zeros((16, 16))
    Ad, Bd, Cd = (torch.from_numpy(t.copy()).to(dev) for t in (A, B, Cm))
    ops.gemm(0, 0, 16, 16, 4, 1.0, Ad, 4, Bd, 16, Cd, 16)
    got = Cd.cpu().numpy()
    ref = A @ B
    if relerr(got, ref) > GEMM_TOL:
        np.set_printoptions(linewidth=200, precision=2)
        say("got\n", got, "\nref\n", ref)
    assert relerr(got, ref) <= GEMM_TOL


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("at,bt", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_f64_shapes(dev, cfg, at, bt):
    ops.set_gemm_config(cfg)
    try:
        for (m, n, k) in [(1, 1, 1), (17, 33, 5), (130, 70, 257), (256, 288, 64), (200, 150, 31)]:
            say("gemm", cfg, at, bt, m, n, k)
            A = RNG.standard_normal((k, m) if at else (m, k))
            B = RNG.standard_normal((n, k) if bt else (k, n))
            C0 = RNG.standard_normal((m, n))
            Ad, Bd = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
            Cd = torch.from_numpy(C0.copy()).to(dev)
            ops.gemm(at, bt, m, n, k, 0.75, Ad, A.shape[1], Bd, B.shape[1], Cd, n)
            ref = C0.copy().ravel()
            oracle.dgemm(at, bt, m, n, k, 0.75, A, A.shape[1], B, B.shape[1], ref, n)
            assert relerr(Cd.cpu().numpy().ravel(), ref) <= GEMM_TOL, (m, n, k)
    finally:
        ops.set_gemm_config(-1)


def test_gemm_leading_dims_and_offsets(dev):
    m, n, k = 45, 38, 29
    lda, ldb, ldc = 33, 41, 50
    A = RNG.standard_normal((m, lda))
    B = RNG.standard_normal((k, ldb))
    C0 = RNG.standard_normal((m, ldc))
    Ad, Bd = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    Cd = torch.from_numpy(C0.copy()).to(dev)
    ops.gemm(0, 0, m, n, k, 1.0, Ad, lda, Bd, ldb, Cd, ldc)
    ref = C0.copy()
    ref[:, :n] += A[:, :k] @ B[:, :n]
    got = Cd.cpu().numpy()
    assert relerr(got, ref) <= GEMM_TOL
    assert np.array_equal(got[:, n:], C0[:, n:])  # untouched padding


@pytest.mark.parametrize("cfg", [0, 1, 2, 3])
def test_gemm_group_f64_vs_oracle(dev, cfg):
    """Grouped launch: 19 problems (two launches of at most 16), ragged
    sizes, odd pitches and 8-byte-offset operands (the scalar load path),
    overwrite and accumulate; every C region checked, padding untouched."""
    import ctypes as C

    from dtcf_amd import blocksparse as bsp
    from dtcf_amd._lib import check, lib

    shapes = [(1, 1, 1), (17, 33, 5), (130, 70, 257), (256, 288, 64), (200, 150, 31),
              (64, 64, 16), (129, 1, 300), (1, 200, 40), (300, 257, 129), (96, 192, 48)]
    shapes = (shapes * 2)[:19]
    rng = np.random.default_rng(11 + cfg)
    for overwrite in (1, 0):
        probs, keep, refs = [], [], []
        for i, (m, n, k) in enumerate(shapes):
            lda, ldb, ldc = k + (i % 3), k + (i % 2), n + 2 * (i % 2)
            off = i % 2  # odd element offset: operands not 16-byte aligned
            A = rng.standard_normal(off + m * lda)
            B = rng.standard_normal(off + n * ldb)
            C0 = rng.standard_normal(m * ldc)
            Ad, Bd = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
            Cd = torch.from_numpy(C0.copy()).to(dev)
            keep += [Ad, Bd, Cd]
            probs.append(bsp.GemmProblem(m, n, k, Ad.data_ptr() + 8 * off, lda,
                                         Bd.data_ptr() + 8 * off, ldb, Cd.data_ptr(), ldc))
            ref = C0.copy() if not overwrite else np.zeros_like(C0)
            if overwrite:
                ref.reshape(m, ldc)[:, n:] = C0.reshape(m, ldc)[:, n:]
            oracle.dgemm(0, 1, m, n, k, -1.25, A[off:], lda, B[off:], ldb, ref, ldc)
            refs.append((Cd, ref, m, n, ldc))
        arr = (bsp.GemmProblem * len(probs))(*probs)
        say("gemm group", cfg, overwrite, len(probs))
        check(lib().dtcf_gemm_group_f64(C.cast(arr, C.c_void_p), len(probs), -1.25, overwrite, cfg,
                                        C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        torch.cuda.synchronize()
        for i, (Cd, ref, m, n, ldc) in enumerate(refs):
            got = Cd.cpu().numpy()
            assert relerr(got, ref) <= GEMM_TOL, (cfg, overwrite, i, shapes[i])
            assert np.array_equal(got.reshape(m, ldc)[:, n:], ref.reshape(m, ldc)[:, n:])


@pytest.mark.parametrize("cfg", [0, 3])
@pytest.mark.parametrize("case", ["tail", "underfilled"])
def test_gemm_group_tail_split(dev, cfg, case):
    """Grouped launches whose last tiles are cut into K slices: 'tail' has
    549 (128x128) / 2188 (64x64) tiles on 512 slots, so the last partial
    round's tiles (all of the last problem's and some of the one before it)
    are split; 'underfilled' has fewer tiles than half the slots, so every
    tile is split.  Accumulate and overwrite, ragged edges, against numpy."""
    import ctypes as C

    from dtcf_amd import blocksparse as bsp
    from dtcf_amd._lib import check, lib

    shapes = ([(2000, 2000, 520), (600, 600, 1040), (2000, 2000, 520), (300, 500, 520)]
              if case == "tail" else [(500, 500, 600), (450, 510, 640), (500, 300, 600), (130, 70, 520)])
    rng = np.random.default_rng(77 + cfg)
    for overwrite in (0, 1):
        probs, keep, refs = [], [], []
        for m, n, k in shapes:
            A = rng.standard_normal((m, k))
            B = rng.standard_normal((n, k))
            C0 = rng.standard_normal((m, n))
            Ad, Bd = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
            Cd = torch.from_numpy(C0.copy()).to(dev)
            keep += [Ad, Bd, Cd]
            probs.append(bsp.GemmProblem(m, n, k, Ad.data_ptr(), k, Bd.data_ptr(), k, Cd.data_ptr(), n))
            refs.append((Cd, (0 if overwrite else C0) + 0.75 * (A @ B.T)))
        arr = (bsp.GemmProblem * len(probs))(*probs)
        say("gemm group split", case, cfg, overwrite)
        check(lib().dtcf_gemm_group_f64(C.cast(arr, C.c_void_p), len(probs), 0.75, overwrite, cfg,
                                        C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        torch.cuda.synchronize()
        for i, (Cd, ref) in enumerate(refs):
            assert relerr(Cd.cpu().numpy(), ref) <= GEMM_TOL, (case, cfg, overwrite, i)


@pytest.mark.parametrize("cfg", [0, 3])
@pytest.mark.parametrize("transa,transb", [(0, 1), (1, 1), (0, 0), (1, 0)])
def test_gemm_group_ex_layouts(dev, transa, transb, cfg):
    """dtcf_gemm_group_ex_f64: A as m x k or k x m, B as n x k or k x n, odd
    pitches and extents, overwrite and accumulate; the first group has 549
    / 2188 tiles (tail split), the second a few small problems."""
    import ctypes as C

    from dtcf_amd import blocksparse as bsp
    from dtcf_amd._lib import check, lib

    rng = np.random.default_rng(5 + 2 * transa + transb + cfg)
    groups = [[(2000, 2000, 520), (600, 600, 1040), (2000, 2000, 520), (300, 500, 520)],
              [(17, 33, 5), (130, 71, 257), (201, 150, 31), (64, 64, 16), (1, 9, 3)]]
    for gi, shapes in enumerate(groups):
        for overwrite in (0, 1):
            probs, keep, refs = [], [], []
            for i, (m, n, k) in enumerate(shapes):
                A = rng.standard_normal((k, m) if transa else (m, k))
                B = rng.standard_normal((n, k) if transb else (k, n))
                lda = A.shape[1] + i % 2
                ldb = B.shape[1] + (i + 1) % 2
                Ap = np.zeros((A.shape[0], lda))
                Ap[:, :A.shape[1]] = A
                Bp = np.zeros((B.shape[0], ldb))
                Bp[:, :B.shape[1]] = B
                C0 = rng.standard_normal((m, n))
                Ad, Bd = torch.from_numpy(Ap).to(dev), torch.from_numpy(Bp).to(dev)
                Cd = torch.from_numpy(C0.copy()).to(dev)
                keep += [Ad, Bd, Cd]
                probs.append(bsp.GemmProblem(m, n, k, Ad.data_ptr(), lda, Bd.data_ptr(), ldb,
                                             Cd.data_ptr(), n))
                opa = A.T if transa else A
                opb = B.T if transb else B
                refs.append((Cd, (0 if overwrite else C0) - 1.5 * (opa @ opb)))
            arr = (bsp.GemmProblem * len(probs))(*probs)
            say("gemm group ex", transa, transb, cfg, gi, overwrite)
            check(lib().dtcf_gemm_group_ex_f64(C.cast(arr, C.c_void_p), len(probs), transa, transb,
                                               -1.5, overwrite, cfg,
                                               C.c_void_p(torch.cuda.current_stream().cuda_stream)))
            torch.cuda.synchronize()
            for i, (Cd, ref) in enumerate(refs):
                assert relerr(Cd.cpu().numpy(), ref) <= GEMM_TOL, (gi, overwrite, i)


@pytest.mark.parametrize("orient", ["col", "row"])
@pytest.mark.parametrize("rows,k,layout", [(5000, 9216, "pairs"), (37, 130, "pairs"), (300, 4, "pairs"),
                                           (700, 20482, "pairs"), (900, 515, "pairs"),
                                           (800, 256, "rows"), (5000, 9216, "rows"), (130, 7, "rows"),
                                           (2, 40000, "rows"), (37, 130, "rows")])
@pytest.mark.parametrize("cview", [False, True])
def test_gemm_view_matrix_vector(dev, orient, rows, k, layout, cview):
    """Matrix-vector products through the view GEMM (n = 1: C[:, 0] = alpha
    A x; m = 1: C[0, :] = alpha x B^T).  The long operand's rows are
    scrambled (a row offset table), its column pairs adjacent ('pairs': the
    streaming path when k is even and at most 20480; k = 20482 and 515 take
    the GEMM) or its rows adjacent ('rows'); x is read through a
    strided view; C is dense or a scrambled view; accumulate into C, against
    numpy.  Row pairs adjacent with an even row count take the row-pair
    streaming kernel (any k); an odd row count takes the GEMM."""
    rng = np.random.default_rng(rows + k)
    Mh = rng.standard_normal((rows, k))
    xh = rng.standard_normal(k)
    perm = rng.permutation(rows)
    if layout == "pairs":        # row r of the view is stored row perm[r], row-major
        Mbuf = Mh[np.argsort(perm)].copy()
        mv = ops.View(rows, k, perm.astype(np.int64) * k, np.arange(k, dtype=np.int64))
    else:                         # column-major storage: rows adjacent
        Mbuf = Mh.T.copy()
        mv = ops.View(rows, k, np.arange(rows, dtype=np.int64), np.arange(k, dtype=np.int64) * rows)
    xbuf = np.zeros(3 * k)
    xbuf[::3] = xh                # x(c) at 3c
    xv = ops.View(1, k, np.zeros(1, dtype=np.int64), 3 * np.arange(k, dtype=np.int64))
    Md, xd = torch.from_numpy(Mbuf.ravel()).to(dev), torch.from_numpy(xbuf).to(dev)
    C0 = rng.standard_normal(rows)
    if orient == "col":           # C rows x 1
        m, n, A, av, B, bv = rows, 1, Md, mv, xd, xv
    else:                         # C 1 x rows
        m, n, A, av, B, bv = 1, rows, xd, xv, Md, mv
    cperm = rng.permutation(rows)
    if cview:
        Cbuf = np.zeros(rows)
        Cbuf[cperm] = C0
        cv = (ops.View(rows, 1, cperm.astype(np.int64), np.zeros(1, dtype=np.int64)) if orient == "col"
              else ops.View(1, rows, np.zeros(1, dtype=np.int64), cperm.astype(np.int64)))
    else:
        Cbuf, cv = C0.copy(), None
    Cd = torch.from_numpy(Cbuf).to(dev)
    say("gemv", orient, rows, k, layout, cview)
    ops.gemm_view(m, n, k, -0.5, A, av, B, bv, Cd, cv, n)
    torch.cuda.synchronize()
    got = Cd.cpu().numpy()
    if cview:
        got = got[cperm]
    ref = C0 - 0.5 * (Mh @ xh)
    assert relerr(got, ref) <= GEMM_TOL


def _random_layout(rng, rows, cols, dev):
    """A rows x cols operand stored in a buffer with spare room, in one of
    four layouts: row-major, column-major, row-major with scrambled rows, or
    element offsets with a stride of 3 (no adjacent pairs).  Returns (dense
    host matrix, device buffer, View)."""
    M = rng.standard_normal((rows, cols))
    kind = rng.integers(4)
    if kind == 0:
        buf, ro, co = M.ravel().copy(), np.arange(rows) * cols, np.arange(cols)
    elif kind == 1:
        buf, ro, co = M.T.ravel().copy(), np.arange(rows), np.arange(cols) * rows
    elif kind == 2:
        perm = rng.permutation(rows)
        buf = M[np.argsort(perm)].ravel().copy()
        ro, co = perm * cols, np.arange(cols)
    else:
        buf = np.zeros(3 * rows * cols)
        ro, co = np.arange(rows) * 3 * cols, np.arange(cols) * 3
        for r in range(rows):
            buf[ro[r] + co] = M[r]
    return M, torch.from_numpy(buf).to(dev), ops.View(rows, cols, ro.astype(np.int64), co.astype(np.int64))


def test_gemm_view_random_shapes(dev):
    """60 seeded random view GEMMs C (+)= alpha A B^T, m, n, k from 1 to
    700 (ones included, so matrix-vector and outer products), every operand
    in a random layout, C dense (with a pitch) or a scrambled view, under
    the automatic tile choice (64 x 64 or 128 x 128, split-K in its three
    forms, the matrix-vector path), against numpy."""
    rng = np.random.default_rng(2024)
    sizes = [1, 2, 3, 17, 64, 65, 127, 128, 130, 255, 256, 700]
    for case in range(60):
        m, n, k = (int(rng.choice(sizes)) for _ in range(3))
        if case % 5 == 0:
            k = int(rng.integers(1025, 3000))      # longer K: 128 x 128 tiles, K splits
        Ah, Ad, av = _random_layout(rng, m, k, dev)
        Bh, Bd, bv = _random_layout(rng, n, k, dev)
        C0 = rng.standard_normal((m, n))
        alpha = float(rng.choice([1.0, -0.75, 2.5]))
        if rng.integers(2):
            perm = rng.permutation(m * n)
            Cbuf = np.zeros(m * n)
            Cbuf[perm] = C0.ravel()
            cv = ops.View(m, n, (perm.reshape(m, n)[:, 0]).astype(np.int64),
                          np.zeros(n, dtype=np.int64)) if n == 1 else None
            if cv is None:     # a scrambled-row view of a row-major C
                rp = rng.permutation(m)
                Cbuf = C0[np.argsort(rp)].ravel().copy()
                cv = ops.View(m, n, (rp * n).astype(np.int64), np.arange(n, dtype=np.int64))
                unscr = lambda x, rp=rp, m=m, n=n: x.reshape(m, n)[rp]
            else:
                unscr = lambda x, perm=perm, m=m, n=n: x[perm].reshape(m, n)
            ldc = n
        else:
            ldc = n + int(rng.integers(0, 3))
            Cbuf = np.zeros(m * ldc)
            Cbuf.reshape(m, ldc)[:, :n] = C0
            cv = None
            unscr = lambda x, m=m, n=n, ldc=ldc: x.reshape(m, ldc)[:, :n]
        Cd = torch.from_numpy(Cbuf).to(dev)
        say("random gemm", case, m, n, k, cv is not None)
        ops.gemm_view(m, n, k, alpha, Ad, av, Bd, bv, Cd, cv, ldc)
        torch.cuda.synchronize()
        got = unscr(Cd.cpu().numpy())
        ref = C0 + alpha * (Ah @ Bh.T)
        assert relerr(got, ref) <= GEMM_TOL, (case, m, n, k)


# (m, n, k, transa, transb): thin in n (the long operand A) or in m (B,
# C written transposed); long operand with k pairs adjacent (A untransposed,
# B transposed) or row pairs adjacent, odd long sides (row pairs: the tile
# path instead), widths 1-48, K from 8 to 200 (the largest at width 48 that
# fits LDS), wide thin sides (49-192, K <= 48: several groups of 48), and
# shapes the thin kernel declines (width 49 with K 64, width 193, K not a
# multiple of 8, K too long).
THIN_SHAPES = [(20000, 48, 192, 0, 0), (20000, 48, 192, 1, 1), (3001, 17, 8, 0, 1), (1000, 16, 48, 1, 0),
               (1001, 1, 16, 1, 0), (48, 5000, 200, 0, 0), (40, 3001, 64, 0, 1), (17000, 32, 96, 1, 1),
               (64, 40, 8, 0, 0), (5, 3, 24, 0, 1), (2, 7000, 48, 1, 1), (4000, 49, 64, 0, 0),
               (3000, 40, 12, 0, 0), (3000, 40, 2048, 0, 1), (192, 5000, 48, 0, 0), (3000, 100, 16, 0, 1),
               (150, 4001, 40, 1, 1), (2999, 192, 48, 1, 0), (4000, 193, 48, 0, 0), (64, 3000, 8, 1, 0)]


@pytest.mark.parametrize("cfg", [6, -1])
def test_gemm_thin_dense(dev, cfg):
    """Thin products through the dense GEMM, the streaming kernel forced
    (cfg 6: every shape it takes, the rest on 64 x 64 tiles) or by the
    automatic choice (long side >= 16384), C accumulated, against numpy."""
    rng = np.random.default_rng(77)
    ops.set_gemm_config(cfg)
    try:
        for (m, n, k, at, bt) in THIN_SHAPES:
            say("thin dense", cfg, m, n, k, at, bt)
            A = rng.standard_normal((k, m) if at else (m, k))
            B = rng.standard_normal((n, k) if bt else (k, n))
            C0 = rng.standard_normal((m, n))
            Cd = torch.from_numpy(C0.copy()).to(dev)
            ops.gemm(at, bt, m, n, k, 0.75, torch.from_numpy(A).to(dev), A.shape[1],
                     torch.from_numpy(B).to(dev), B.shape[1], Cd, n)
            ref = C0 + 0.75 * ((A.T if at else A) @ (B.T if bt else B))
            assert relerr(Cd.cpu().numpy(), ref) <= GEMM_TOL, (m, n, k, at, bt)
    finally:
        ops.set_gemm_config(-1)


def test_gemm_thin_views(dev):
    """Thin products through the view GEMM with the streaming kernel forced:
    operands in random layouts (row-major, column-major, scrambled rows, no
    adjacent pairs), C dense with a pitch or a scrambled-row view, thin in n
    or in m, against numpy."""
    rng = np.random.default_rng(4242)
    shapes = [(3000, 48, 192), (2998, 33, 48), (40, 2500, 200), (1999, 1, 8), (16, 4000, 64), (4001, 20, 8),
              (192, 3000, 48), (2500, 100, 24)]
    ops.set_gemm_config(6)
    try:
        for case in range(32):
            m, n, k = shapes[case % len(shapes)]
            Ah, Ad, av = _random_layout(rng, m, k, dev)
            Bh, Bd, bv = _random_layout(rng, n, k, dev)
            C0 = rng.standard_normal((m, n))
            if case % 2:
                rp = rng.permutation(m)
                Cd = torch.from_numpy(C0[np.argsort(rp)].ravel().copy()).to(dev)
                cv = ops.View(m, n, (rp * n).astype(np.int64), np.arange(n, dtype=np.int64))
                ldc, unscr = n, (lambda x, rp=rp, m=m, n=n: x.reshape(m, n)[rp])
            else:
                ldc = n + 1
                Cbuf = np.zeros(m * ldc)
                Cbuf.reshape(m, ldc)[:, :n] = C0
                Cd, cv = torch.from_numpy(Cbuf).to(dev), None
                unscr = lambda x, m=m, n=n, ldc=ldc: x.reshape(m, ldc)[:, :n]
            say("thin view", case, m, n, k, cv is not None)
            ops.gemm_view(m, n, k, -1.25, Ad, av, Bd, bv, Cd, cv, ldc)
            torch.cuda.synchronize()
            ref = C0 - 1.25 * (Ah @ Bh.T)
            assert relerr(unscr(Cd.cpu().numpy()), ref) <= GEMM_TOL, (case, m, n, k)
    finally:
        ops.set_gemm_config(-1)


@pytest.mark.parametrize("m,n,k", [(2200, 2100, 520), (2049, 2300, 1040)])
def test_gemm_one_partial_round_split(dev, m, n, k):
    """More tiles than half the CU slots but at most all of them: every tile
    is cut into K slices filling whole rounds (e.g. 306 tiles -> 3 slices);
    accumulate into an existing C, edge tiles included, against numpy."""
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((m, k))
    B = rng.standard_normal((n, k))
    C0 = rng.standard_normal((m, n))
    Cd = torch.from_numpy(C0.copy()).to(dev)
    ops.gemm(0, 1, m, n, k, 0.5, torch.from_numpy(A).to(dev), k, torch.from_numpy(B).to(dev), k,
             Cd, n)
    ref = C0 + 0.5 * (A @ B.T)
    assert relerr(Cd.cpu().numpy(), ref) <= GEMM_TOL


def test_gemm_repeat_is_bitwise_deterministic(dev):
    m = n = k = 384
    A = torch.from_numpy(RNG.standard_normal((m, k))).to(dev)
    B = torch.from_numpy(RNG.standard_normal((n, k))).to(dev)
    outs = []
    for _ in range(3):
        Cd = torch.zeros((m, n), dtype=torch.float64, device=dev)
        ops.gemm(0, 1, m, n, k, 1.0, A, k, B, k, Cd, n)
        outs.append(Cd.cpu())
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[1], outs[2])


def _random_view(rows, cols, rpb, cpb, kind, rng):
    """A view over a buffer of random tiles with a random intra-tile permutation."""
    nrb, ncb = rows // rpb, cols // cpb
    bs = rpb * cpb
    if kind == "col":       # tile row-major [rpb][cpb]
        r0 = np.arange(rpb) * cpb
        c0 = np.arange(cpb)
    elif kind == "row":     # tile stored transposed [cpb][rpb]
        r0 = np.arange(rpb)
        c0 = np.arange(cpb) * rpb
    else:                   # neither direction contiguous: strided in both
        r0 = np.arange(rpb) * 2 * cpb
        c0 = np.arange(cpb) * 2
        bs = 2 * rpb * cpb
    order = rng.permutation(nrb * ncb)
    blk = order * bs
    rofs = np.tile(r0, nrb)
    cofs = np.tile(c0, ncb)
    return rofs, cofs, blk, bs * nrb * ncb


def _view_dense(buf, rofs, cofs, blk, rpb, cpb, ncb):
    rows, cols = len(rofs), len(cofs)
    R = np.arange(rows) // rpb
    Cb = np.arange(cols) // cpb
    idx = rofs[:, None] + cofs[None, :] + blk[R[:, None] * ncb + Cb[None, :]]
    return buf[idx]


@pytest.mark.parametrize("ka,kb,kc", [("col", "col", "col"), ("col", "row", "col"),
                                      ("row", "col", "row"), ("row", "row", "col"),
                                      ("str", "col", "col")])
def test_gemm_view_matches_dense(dev, ka, kb, kc):
    rng = np.random.default_rng(11)
    m, n, k = 96, 80, 72
    rpa, cpa, rpbb, cpbb = 12, 24, 20, 24
    ra, ca, ba, sa = _random_view(m, k, rpa, cpa, ka, rng)
    rb, cb, bb, sb = _random_view(n, k, rpbb, cpbb, kb, rng)
    rc, cc, bc, sc = _random_view(m, n, rpa, rpbb, kc, rng)
    Abuf, Bbuf, Cbuf = rng.standard_normal(sa), rng.standard_normal(sb), rng.standard_normal(sc)
    Ad = _view_dense(Abuf, ra, ca, ba, rpa, cpa, k // cpa)
    Bd = _view_dense(Bbuf, rb, cb, bb, rpbb, cpbb, k // cpbb)
    Cd0 = _view_dense(Cbuf, rc, cc, bc, rpa, rpbb, n // rpbb)
    av = ops.View(m, k, ra, ca, ba, rpa, cpa)
    bv = ops.View(n, k, rb, cb, bb, rpbb, cpbb)
    cv = ops.View(m, n, rc, cc, bc, rpa, rpbb)
    assert av.span <= sa and bv.span <= sb and cv.span <= sc
    Ag, Bg, Cg = (torch.from_numpy(x.copy()).to(dev) for x in (Abuf, Bbuf, Cbuf))
    ops.gemm_view(m, n, k, -1.25, Ag, av, Bg, bv, Cg, cv)
    got = _view_dense(Cg.cpu().numpy(), rc, cc, bc, rpa, rpbb, n // rpbb)
    ref = Cd0 - 1.25 * Ad @ Bd.T
    assert relerr(got, ref) <= GEMM_TOL
    # gather / scatter round trip is bit-exact
    big = torch.empty((m, k), dtype=torch.float64, device=dev)
    ops.view_gather(Ag, av, big, k)
    assert np.array_equal(big.cpu().numpy(), Ad)
    back = torch.zeros_like(Ag)
    ops.view_scatter(big, k, av, back)
    touched = np.zeros(sa, bool)
    idx = (ra[:, None] + ca[None, :] + ba[(np.arange(m) // rpa)[:, None] * (k // cpa)
                                          + (np.arange(k) // cpa)[None, :]])
    touched[idx.ravel()] = True
    assert np.array_equal(back.cpu().numpy()[touched], Abuf[touched])


def _c2_like(o, v, vg, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((o, o, v, v)) * 0.05
    B = rng.standard_normal((o, o, v, v)) * 0.05
    Cm = rng.standard_normal((o, o, v, v)) * 0.05
    return A, B, Cm


@pytest.mark.parametrize("mode", [rrr.LITERAL, rrr.PACKED, rrr.FUSED])
def test_local_step_ring_modes(dev, mode):
    # C[i,j,a,b] += sum_{k,c} A[i,k,a,c] B[k,j,c,b]  (the ring term of config 2, small)
    o, v = 8, 24
    vg = [2, 2, 3, 3]
    A, B, Cm = _c2_like(o, v, vg, 3)
    tA = rrr.TileTensor.from_dense("cccc", A, vg, dev)
    tB = rrr.TileTensor.from_dense("cccc", B, vg, dev)
    tC = rrr.TileTensor.from_dense("cccc", Cm, vg, dev)
    say("step mode", mode)
    st = rrr.LocalStep(tA, "i,k,a,c", tB, "k,j,c,b", tC, "i,j,a,b", mode=mode)
    assert st.big_mnk == (o * v, o * v, o * v)
    st.run(1.0)
    torch.cuda.synchronize()
    got = rrr.dense_from_tiles(tC.tiles.cpu().numpy(), Cm.shape, vg, tC.addrs)
    ref = Cm + np.einsum("ikac,kjcb->ijab", A, B)
    assert relerr(got, ref) <= GEMM_TOL


def test_local_step_ladder_and_transposed_output(dev):
    # C1 shape (scaled down): R[a,b,i,j] += 0.5 * sum_{cd} T[c,d,i,j] V[a,b,c,d]
    o, v = 4, 8
    rng = np.random.default_rng(5)
    T = rng.standard_normal((v, v, o, o))
    V = rng.standard_normal((v, v, v, v))
    R = rng.standard_normal((v, v, o, o))
    vgT, vgV, vgR = [2, 2, 2, 1], [2, 2, 2, 2], [2, 2, 2, 1]
    for mode in (rrr.LITERAL, rrr.PACKED, rrr.FUSED):
        tT = rrr.TileTensor.from_dense("cccc", T, vgT, dev)
        tV = rrr.TileTensor.from_dense("cccc", V, vgV, dev)
        tR = rrr.TileTensor.from_dense("cccc", R, vgR, dev)
        st = rrr.LocalStep(tT, "c,d,i,j", tV, "a,b,c,d", tR, "a,b,i,j", mode=mode)
        st.run(0.5)
        got = rrr.dense_from_tiles(tR.tiles.cpu().numpy(), R.shape, vgR, tR.addrs)
        ref = R + 0.5 * np.einsum("cdij,abcd->abij", T, V)
        assert relerr(got, ref) <= GEMM_TOL, mode


def _bounced(sym, full, vg, dev):
    """Tiles of a tensor symmetric in its first two dims: only tiles with
    addr0 >= addr1 are stored; a bounced copy arrives under the swapped
    address carrying the stored tile's data, and the plan must swap its dims
    back without a sign (get_symm_permutation, contraction.cc:132-160)."""
    stored = rrr.tile_addresses(sym, vg)
    br = [full.shape[d] // vg[d] for d in range(4)]
    addrs, datas = [], []
    for a in stored:
        sl = tuple(slice(a[d] * br[d], (a[d] + 1) * br[d]) for d in range(4))
        blk = full[sl].ravel()
        addrs.append(list(a))
        datas.append(blk)
        if a[0] > a[1]:
            addrs.append([a[1], a[0], a[2], a[3]])
            datas.append(blk)  # same data under the bounced address
    return rrr.TileTensor(sym, full.shape, vg, dev,
                          data=torch.from_numpy(np.concatenate(datas)).to(dev), addrs=addrs)


@pytest.mark.parametrize("mode", [rrr.LITERAL, rrr.PACKED, rrr.FUSED])
def test_local_step_symmetric_bounce(dev, mode):
    # A is permutationally symmetric in (i,k) ('aacc'): two bounce classes.
    # LITERAL permutes class by class into a buffer and gathers; PACKED and
    # FUSED pack each class straight into the big matrix.
    o, v = 6, 8
    vg = [3, 3, 2, 2]
    rng = np.random.default_rng(9)
    X = rng.standard_normal((o, o, v, v))
    full = X + X.transpose(1, 0, 2, 3)
    tA = _bounced("aacc", full, vg, dev)
    B = rng.standard_normal((o, o, v, v))
    tB = rrr.TileTensor.from_dense("cccc", B, vg, dev)
    Cm = rng.standard_normal((o, o, v, v))
    tC = rrr.TileTensor.from_dense("cccc", Cm, vg, dev)
    say("symmetric bounce, mode", mode)
    st = rrr.LocalStep(tA, "i,k,a,c", tB, "k,j,c,b", tC, "i,j,a,b", mode=mode)
    st.run(1.0)
    got = rrr.dense_from_tiles(tC.tiles.cpu().numpy(), Cm.shape, vg, tC.addrs)
    ref = Cm + np.einsum("ikac,kjcb->ijab", full, B)
    assert relerr(got, ref) <= GEMM_TOL


@pytest.mark.parametrize("which", ["AB", "B"])
@pytest.mark.parametrize("mode", [rrr.LITERAL, rrr.PACKED, rrr.FUSED])
def test_local_step_both_operands_symmetric(dev, mode, which):
    # A symmetric in (i,k), B in (k,j) ('aacc'), or B alone, uneven tile
    # extents (o=9 in 3 tiles, v=10 in 2) and the contraction index k split
    # over row and column roles of the two operands: the bounced operands
    # are packed by class, the others and C read through their views.
    o, v = 9, 10
    vg = [3, 3, 2, 2]
    rng = np.random.default_rng(21)
    X = rng.standard_normal((o, o, v, v))
    fa = X + X.transpose(1, 0, 2, 3) if "A" in which else X
    Y = rng.standard_normal((o, o, v, v))
    fb = Y + Y.transpose(1, 0, 2, 3)
    tA = _bounced("aacc", fa, vg, dev) if "A" in which else rrr.TileTensor.from_dense("cccc", fa, vg, dev)
    tB = _bounced("aacc", fb, vg, dev)
    Cm = rng.standard_normal((o, o, v, v))
    tC = rrr.TileTensor.from_dense("cccc", Cm, vg, dev)
    say("symmetric", which, "mode", mode)
    st = rrr.LocalStep(tA, "i,k,a,c", tB, "k,j,c,b", tC, "i,j,a,b", mode=mode)
    for _ in range(2):  # a second run re-packs over the first run's big matrices
        st.run(0.5)
    got = rrr.dense_from_tiles(tC.tiles.cpu().numpy(), Cm.shape, vg, tC.addrs)
    ref = Cm + np.einsum("ikac,kjcb->ijab", fa, fb)
    assert relerr(got, ref) <= GEMM_TOL


@pytest.mark.parametrize("o,v", [(6, 10), (5, 7), (12, 70)])
def test_reductions_vs_oracle(dev, o, v):
    # (6, 10) and (12, 70): 16-byte paths, several 64 x 64 energy tiles;
    # (5, 7): odd o*o and v*v, the element paths
    eo = -2.0 + 1.5 * RNG.random(o)
    ev = 0.2 + 2.8 * RNG.random(v)
    R = RNG.standard_normal(v * v * o * o) * 0.05
    Vm = RNG.standard_normal(o * o * v * v) * 0.05
    Rd = torch.from_numpy(R).to(dev)
    T = ops.denom_update(Rd, torch.from_numpy(eo).to(dev), torch.from_numpy(ev).to(dev))
    Tref = oracle.denom_update(R, eo, ev)
    assert np.array_equal(T.cpu().numpy(), Tref)  # IEEE division, same operation order
    E = ops.energy(torch.from_numpy(Vm).to(dev), T, o, v).item()
    Eref = oracle.energy(Vm, Tref, o, v)
    assert abs(E - Eref) <= 1e-13 * abs(Eref)
    n2 = ops.norm2(Rd).item()
    assert abs(n2 - oracle.norm2(R)) <= 1e-13 * oracle.norm2(R)
    E2 = ops.energy(torch.from_numpy(Vm).to(dev), T, o, v).item()
    assert E2 == E  # deterministic
