"""Pins of the CPU oracle (CPU only).

Each oracle function is checked against something independent of its own
formula: the reference's own DLTC kernels compiled from source
(oracle/_ref), numpy library routines (transpose, matmul, einsum,
ravel_multi_index), libc's drand48, closed forms, and brute force.
"""
import ctypes
import itertools

import numpy as np
import pytest

import oracle

RNG = np.random.default_rng(1234)


def brute_rrr(x, pmap, w):
    out_src = [0] * len(pmap)
    for i, p in enumerate(pmap):
        out_src[p] = i
    return np.ascontiguousarray(np.transpose(x.reshape(w), out_src)).ravel()


@pytest.mark.parametrize("w", [[3, 5], [4, 6, 5], [2, 3, 4, 5], [4, 8, 4, 12], [3, 2, 5, 4, 3],
                               [2, 3, 2, 4, 3, 2], [12, 12, 48, 48]])
def test_rrr_transpose_all_paths(w):
    x = RNG.standard_normal(int(np.prod(w)))
    perms = list(itertools.permutations(range(len(w))))
    if len(perms) > 40:
        perms = [perms[i] for i in RNG.choice(len(perms), 40, replace=False)] + [tuple(range(len(w)))]
    for pm in perms:
        got = oracle.rrr_transpose(x, pm, w)
        assert np.array_equal(got, brute_rrr(x, pm, w)), pm


def test_rrr_transpose_detects_wrong_map():
    # a planted mistake (inverse map instead of the map) must be visible
    w = [2, 3, 4, 5]
    x = RNG.standard_normal(120)
    pm = [1, 2, 3, 0]
    inv = [0] * 4
    for i, p in enumerate(pm):
        inv[p] = i
    assert not np.array_equal(oracle.rrr_transpose(x, inv, w), brute_rrr(x, pm, w))


def test_rrr_transpose_vector_width_paths():
    # element path with VECTOR_WIDTH 1, 2, 3, 4, 32 (helper.cc:465-473)
    for w in ([5, 7], [4, 6], [6, 9], [8, 12], [64, 32], [3, 64, 96]):
        x = RNG.standard_normal(int(np.prod(w)))
        pm = list(range(len(w)))[::-1]
        assert np.array_equal(oracle.rrr_transpose(x, pm, w), brute_rrr(x, pm, w))


def test_dltc_position():
    x = [3, 7, 1, 5]
    y = [1, 5, 3, 7]
    assert list(oracle.dltc_position(x, y)) == [y.index(v) for v in x]


REF_SHAPES = [(4, 4, 4, 4), (5, 6, 7, 9), (8, 3, 12, 5), (1, 9, 2, 4), (13, 4, 6, 7)]


@pytest.mark.parametrize("kind", ["seq", "omp"])
def test_dltc_kernels_match_reference(kind):
    ref = oracle.ref_transp(kind)
    if ref is None:
        pytest.skip("oracle/_ref not built (needs /root/reference)")
    for shp in REF_SHAPES:
        n = int(np.prod(shp))
        x = RNG.standard_normal(n)
        for name in oracle.DLTC_4D_NAMES:
            out = np.full(n, np.nan)
            ref[name](x, out, *shp)
            s = ["ijkl".index(ch) for ch in name]
            mine = oracle.permute_by_source(x, s, list(shp))
            assert np.array_equal(out, mine), (name, shp)
            # the dispatcher: x = (0,1,2,3) -> y with y[n] = x[s[n]]
            st, y = oracle.dltc_transpose(x, [0, 1, 2, 3], s, list(shp))
            assert st == 0 and np.array_equal(y, out)
    for shp in [(5, 7), (32, 33), (64, 3), (40, 40)]:
        x = RNG.standard_normal(shp[0] * shp[1])
        for name, s in (("ij", [0, 1]), ("ji", [1, 0])):
            out = np.full(x.size, np.nan)
            ref[name](x, out, *shp)
            assert np.array_equal(out, oracle.permute_by_source(x, s, list(shp))), (name, shp)


def test_dltc_identity_and_errors():
    x = RNG.standard_normal(24)
    st, y = oracle.dltc_transpose(x, [0, 1, 2, 3], [0, 1, 2, 3], [2, 3, 2, 2])
    assert st == 0 and np.array_equal(y, x)
    # the reference's literal behaviour leaves the output untouched (Q2)
    sentinel = np.full(24, -7.0)
    st, y = oracle.dltc_transpose(x, [0, 1, 2, 3], [0, 1, 2, 3], [2, 3, 2, 2], q2_literal=True,
                                  out=sentinel)
    assert st == 0 and np.all(y == -7.0)
    st, _ = oracle.dltc_transpose(x[:6], [0, 1, 2], [2, 1, 0], [1, 2, 3])
    assert st == -1


def test_block_grid_closed_form():
    grid = [3, 4, 5, 2]
    for r in range(int(np.prod(grid))):
        a = oracle.block_addr(r, grid)
        assert list(a) == list(np.unravel_index(r, grid))
        assert oracle.block_rank(a, grid) == np.ravel_multi_index(tuple(a), grid)


def test_symm_permutation_cases():
    # 'aabb': groups 0,0,1,1
    sg = [0, 0, 1, 1]
    assert list(oracle.symm_permutation([2, 1, 3, 0], sg)) == [0, 1, 2, 3]
    assert list(oracle.symm_permutation([1, 2, 3, 0], sg)) == [1, 0, 2, 3]
    assert list(oracle.symm_permutation([1, 2, 0, 3], sg)) == [1, 0, 3, 2]
    # non-symmetric groups never swap
    assert list(oracle.symm_permutation([0, 3, 1, 2], [3, 3, 3, 3])) == [0, 1, 2, 3]
    # group of three: each i swaps with its first larger j and breaks
    # (contraction.cc:146-155); for addresses (0,1,2) the later swap of i=1
    # overwrites sym[1], leaving [1, 2, 1] -- not a permutation, so the
    # reference only handles symmetry groups of two dims.
    assert list(oracle.symm_permutation([0, 1, 2], [0, 0, 0])) == [1, 2, 1]
    assert list(oracle.compose_p_maps([1, 0, 2, 3], [0, 2, 1, 3])) == [2, 0, 1, 3]


def test_big_matrix_order_and_round_trip():
    vgrid = [3, 2, 4, 2]
    pmap = [0, 2, 1, 3]
    addrs = np.array(list(itertools.product(*[range(v) for v in vgrid])), dtype=np.int32)
    RNG.shuffle(addrs)
    perm_addr = np.zeros_like(addrs)
    for d in range(4):
        perm_addr[:, pmap[d]] = addrs[:, d]
    order, nr, nc = oracle.big_matrix_order(vgrid, pmap, perm_addr, 2)
    grid = [0] * 4
    for d in range(4):
        grid[pmap[d]] = vgrid[d]
    # brute force: ascending (row rank, col rank)
    keys = []
    for t, pa in enumerate(perm_addr):
        row = np.ravel_multi_index((pa[0], pa[1], 0, 0), grid)
        col = np.ravel_multi_index((0, 0, pa[2], pa[3]), grid)
        keys.append((row, col, t))
    keys.sort()
    assert [k[2] for k in keys] == list(order)
    assert nr == grid[0] * grid[1] and nc == grid[2] * grid[3]
    rpb, cpb = 3, 5
    tiles = RNG.standard_normal(len(addrs) * rpb * cpb)
    big = oracle.big_matrix_pack(tiles, order, nr, nc, rpb, cpb)
    # closed form: block (r, c) holds tile order[r*nc+c]
    for r in range(nr):
        for c in range(nc):
            t = order[r * nc + c]
            assert np.array_equal(big[r * rpb:(r + 1) * rpb, c * cpb:(c + 1) * cpb],
                                  tiles[t * rpb * cpb:(t + 1) * rpb * cpb].reshape(rpb, cpb))
    back = oracle.big_matrix_unpack(big, order, nr, nc, rpb, cpb, len(addrs))
    assert np.array_equal(back, tiles)   # big_matrix_transpose_test (test.cc:698-724)


def test_big_matrix_ragged_rows_rejected():
    vgrid = [2, 2]
    addrs = np.array([[0, 0], [0, 1], [1, 0]], dtype=np.int32)
    with pytest.raises(ValueError):
        oracle.big_matrix_order(vgrid, [0, 1], addrs, 1)


@pytest.mark.parametrize("at,bt", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_dgemm_vs_numpy(at, bt):
    m, n, k = 7, 5, 9
    A = RNG.standard_normal((k, m) if at else (m, k))
    B = RNG.standard_normal((n, k) if bt else (k, n))
    C0 = RNG.standard_normal((m, n))
    Cm = C0.copy().ravel()
    oracle.dgemm(at, bt, m, n, k, 0.5, A, A.shape[1], B, B.shape[1], Cm, n)
    ref = C0 + 0.5 * (A.T if at else A) @ (B.T if bt else B)
    assert np.allclose(Cm.reshape(m, n), ref, rtol=1e-14, atol=1e-14)


def test_dgemm_blocked_vs_numpy():
    m, n, k = 130, 70, 300
    A = RNG.standard_normal((m, k))
    B = RNG.standard_normal((n, k))
    Cm = np.zeros(m * n)
    oracle.dgemm_nt_blocked(m, n, k, 1.0, A, k, B, k, Cm, n, 2)
    assert np.allclose(Cm.reshape(m, n), A @ B.T, rtol=1e-13, atol=1e-12)


def test_checkers_vs_einsum():
    O, V = 4, 4
    A = RNG.standard_normal(V * O * O * O)
    B = RNG.standard_normal(V * O * O * O)
    sC = oracle.checkM(A, B, O, V)
    ref = np.einsum("akil,bljk->abij", A.reshape(V, O, O, O), B.reshape(V, O, O, O))
    assert np.allclose(sC, ref.ravel(), rtol=1e-14, atol=1e-14)
    A2 = RNG.standard_normal(V * V * O * O)
    s2 = oracle.checkS(A2, O, V)
    assert np.array_equal(s2, np.transpose(A2.reshape(V, V, O, O), (1, 0, 3, 2)).ravel())


def test_drand48_matches_libc():
    libc = ctypes.CDLL("libc.so.6")
    libc.drand48.restype = ctypes.c_double
    libc.srand48.argtypes = [ctypes.c_long]
    # srand48(s) sets X = (s << 16) | 0x330E (POSIX); glibc's default state is
    # the same as srand48(0x1234ABCD).
    libc.srand48(0x1234ABCD)
    want = np.array([libc.drand48() for _ in range(100)])
    got = oracle.Drand48().fill(100)
    assert np.array_equal(got, want)


def test_cc_closed_forms():
    o, v = 3, 4
    eo = np.array([-1.5, -1.0, -0.7])
    ev = np.array([0.3, 0.9, 1.7, 2.2])
    R = RNG.standard_normal(v * v * o * o)
    T = oracle.denom_update(R, eo, ev)
    D = (eo[None, None, :, None] + eo[None, None, None, :] - ev[:, None, None, None]
         - ev[None, :, None, None])
    assert np.array_equal(T, (R.reshape(v, v, o, o) / D).ravel())
    Vm = RNG.standard_normal(o * o * v * v)
    E = oracle.energy(Vm, T, o, v)
    Eref = 0.25 * np.einsum("ijab,abij->", Vm.reshape(o, o, v, v), T.reshape(v, v, o, o))
    assert abs(E - Eref) <= 1e-13 * max(1.0, abs(Eref))
    # constant inputs: E = o^2 v^2 / 4
    assert oracle.energy(np.ones(o * o * v * v), np.ones(o * o * v * v), o, v) == o * o * v * v / 4
    assert abs(oracle.norm2(R) - np.linalg.norm(R)) <= 1e-13 * np.linalg.norm(R)


def test_ccd_iteration_vs_loops():
    """The precision-study iteration: ladder and ring by brute-force loops,
    then T_new = R / D and E = 1/4 sum V_oovv T_new element by element."""
    o, v = 2, 3
    rng = np.random.default_rng(12)
    V = rng.standard_normal((v, v, v, v))
    Vvvoo = rng.standard_normal((v, v, o, o))
    W = rng.standard_normal((o, o, v, v))
    T = rng.standard_normal((v, v, o, o))
    eo = np.array([-1.5, -0.8])
    ev = np.array([0.3, 1.1, 2.6])
    lad, ring, Tn, E = oracle.ccd_iteration(V, Vvvoo, W, T, eo, ev)
    L = np.zeros((v, v, o, o))
    Rg = np.zeros((v, v, o, o))
    for a, b, i, j in itertools.product(range(v), range(v), range(o), range(o)):
        for c, d in itertools.product(range(v), range(v)):
            L[a, b, i, j] += 0.5 * V[a, b, c, d] * T[c, d, i, j]
        for k, c in itertools.product(range(o), range(v)):
            Rg[a, b, i, j] += T[a, c, i, k] * W[k, j, c, b]
    np.testing.assert_allclose(lad, L, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(ring, Rg, rtol=1e-13, atol=1e-13)
    Eref = 0.0
    for a, b, i, j in itertools.product(range(v), range(v), range(o), range(o)):
        t = (Vvvoo[a, b, i, j] + L[a, b, i, j] + Rg[a, b, i, j]) / (eo[i] + eo[j] - ev[a] - ev[b])
        assert abs(Tn[a, b, i, j] - t) <= 1e-13 * max(1.0, abs(t))
        Eref += 0.25 * Vvvoo[a, b, i, j] * t
    assert abs(E - Eref) <= 1e-13 * max(1.0, abs(Eref))


def _uhf_random(Oa, Ob, Va, Vb, seed, residuals=True):
    rng = np.random.default_rng(seed)
    a = {"eoa": -2.0 + 1.5 * rng.random(Oa), "eob": -2.0 + 1.5 * rng.random(Ob),
         "eva": 0.2 + 2.8 * rng.random(Va), "evb": 0.2 + 2.8 * rng.random(Vb),
         "fa_ov": rng.standard_normal((Oa, Va)), "fb_ov": rng.standard_normal((Ob, Vb)),
         "vaa_oovv": rng.standard_normal((Oa, Oa, Va, Va)),
         "vab_oovv": rng.standard_normal((Oa, Ob, Va, Vb)),
         "vbb_oovv": rng.standard_normal((Ob, Ob, Vb, Vb))}
    shapes = {"ta_vo": (Va, Oa), "tb_vo": (Vb, Ob), "taa_vvoo": (Va, Va, Oa, Oa),
              "tab_vvoo": (Va, Vb, Oa, Ob), "tbb_vvoo": (Vb, Vb, Ob, Ob)}
    for n, shp in shapes.items():
        a[n] = 0.1 * rng.standard_normal(shp)
        if residuals:
            a["r" + n[1:]] = rng.standard_normal(shp)
    return a


def test_uhf_energy_vs_spin_orbital_embedding():
    """The UHF block energy (factors 1/4 same spin, 1 mixed, the singles
    products) equals the spin-orbital CCSD energy of the embedded tensors,
    for antisymmetric and for unsymmetric same-spin blocks, before and after
    an update."""
    for anti in (False, True):
        a = _uhf_random(3, 2, 4, 5, 21 + anti)
        if anti:
            for k in ("vaa_oovv", "vbb_oovv", "taa_vvoo", "tbb_vvoo"):
                x = a[k]
                a[k] = 0.5 * (x - x.transpose(1, 0, 2, 3) - x.transpose(0, 1, 3, 2)
                              + x.transpose(1, 0, 3, 2))
        for update in (False, True):
            t, out = oracle.uhf_ccsd_step(a, update)
            e_so = oracle.spin_orbital_energy(a, t)
            assert abs(out[0] - e_so) <= 1e-13 * max(1.0, abs(e_so)), (anti, update)
            assert abs(out[0] - out[1:5].sum()) <= 1e-14 * max(1.0, abs(out[0]))


def test_uhf_update_loops_and_zero_start():
    a = _uhf_random(2, 3, 4, 3, 5)
    t, out = oracle.uhf_ccsd_step(a, True)
    for n, eo, ev in (("ta_vo", a["eoa"], a["eva"]), ("tb_vo", a["eob"], a["evb"])):
        r = a["r" + n[1:]]
        for x, i in itertools.product(range(len(ev)), range(len(eo))):
            assert t[n][x, i] == a[n][x, i] + r[x, i] / (eo[i] - ev[x])
    for n, e1, e2, v1, v2 in (("taa_vvoo", "eoa", "eoa", "eva", "eva"),
                              ("tab_vvoo", "eoa", "eob", "eva", "evb"),
                              ("tbb_vvoo", "eob", "eob", "evb", "evb")):
        r = a["r" + n[1:]]
        for x, y, i, j in itertools.product(*[range(k) for k in r.shape]):
            d = ((a[e1][i] + a[e2][j]) - a[v1][x]) - a[v2][y]
            assert t[n][x, y, i, j] == a[n][x, y, i, j] + r[x, y, i, j] / d
    r1 = np.sqrt(sum(np.sum(a[k] ** 2) for k in ("ra_vo", "rb_vo")))
    r2 = np.sqrt(sum(np.sum(a[k] ** 2) for k in ("raa_vvoo", "rab_vvoo", "rbb_vvoo")))
    dt = np.sqrt(sum(np.sum((t[k] - a[k]) ** 2) for k in t))
    assert abs(out[5] - r1) <= 1e-13 * r1 and abs(out[6] - r2) <= 1e-13 * r2
    assert abs(out[7] - dt) <= 1e-12 * dt
    # zero amplitudes, residuals f_vo and v_vvoo: first-order amplitudes, and
    # the energy is sum f_ia f_ai / D_ai + 1/4 sum |v|^2 / D (same spin) ...
    for n in ("ta_vo", "tb_vo", "taa_vvoo", "tab_vvoo", "tbb_vvoo"):
        a[n] = np.zeros_like(a[n])
    a["ra_vo"], a["rb_vo"] = a["fa_ov"].T.copy(), a["fb_ov"].T.copy()
    a["raa_vvoo"] = a["vaa_oovv"].transpose(2, 3, 0, 1).copy()
    a["rab_vvoo"] = a["vab_oovv"].transpose(2, 3, 0, 1).copy()
    a["rbb_vvoo"] = a["vbb_oovv"].transpose(2, 3, 0, 1).copy()
    t, out = oracle.uhf_ccsd_step(a, True)
    Da = a["eoa"][None, :] - a["eva"][:, None]
    Db = a["eob"][None, :] - a["evb"][:, None]
    e1 = np.sum(a["fa_ov"].T ** 2 / Da) + np.sum(a["fb_ov"].T ** 2 / Db)
    assert abs(out[1] - e1) <= 1e-13 * abs(e1)
    ta, tb = a["fa_ov"].T / Da, a["fb_ov"].T / Db
    Dab = (a["eoa"][None, None, :, None] + a["eob"][None, None, None, :]
           - a["eva"][:, None, None, None] - a["evb"][None, :, None, None])
    eab = np.sum(a["vab_oovv"].transpose(2, 3, 0, 1) ** 2 / Dab) + \
        np.einsum("ijab,ai,bj->", a["vab_oovv"], ta, tb)
    assert abs(out[3] - eab) <= 1e-12 * abs(eab)


def test_permutation_maps_c1_c2():
    # C2: contract("i,k,a,c", "k,j,c,b", "i,j,a,b")
    br = dict(i=12, j=12, k=12, a=48, b=48, c=48)
    pm = oracle.permutation_maps(list("ikac"), list("kjcb"), list("ijab"), br)
    assert pm["p_map_A"] == [0, 2, 1, 3]
    assert pm["p_map_B"] == [2, 0, 3, 1]
    assert pm["p_map_C"] == [0, 2, 1, 3]
    assert pm["inv_p_map_C"] == [0, 2, 1, 3]
    assert pm["n_a"] == 12 * 48 and pm["n_b"] == 12 * 48 and pm["n_k"] == 12 * 48
    # C1 ladder with A = T[c,d,i,j], B = V[a,b,c,d], C = R[a,b,i,j]
    br = dict(a=32, b=32, c=32, d=32, i=8, j=8)
    pm = oracle.permutation_maps(list("cdij"), list("abcd"), list("abij"), br)
    assert pm["p_map_A"] == [2, 3, 0, 1]
    assert pm["p_map_B"] == [0, 1, 2, 3]
    assert pm["p_map_C"] == [2, 3, 0, 1]


def test_dltc_index_split_ladder():
    # ccd-t2 m[8]: R[p3,p4,h1,h2] += T[p5,p6,h1,h2] * V[p3,p4,p5,p6]
    h = {f"h{i}": i - 1 for i in range(1, 13)}
    p = {f"p{i}": 11 + i for i in range(1, 13)}
    ida = [p["p5"], p["p6"], h["h1"], h["h2"]]
    idb = [p["p3"], p["p4"], p["p5"], p["p6"]]
    sp = oracle.dltc_index_split(ida, idb)
    assert sp["m"] == [h["h1"], h["h2"]]
    assert sp["n"] == [p["p3"], p["p4"]]
    assert sp["k"] == [p["p5"], p["p6"]]
    # position(at, ida) gives the transpose (klij)
    s = oracle.dltc_position(sp["at"], ida)
    assert list(s) == [2, 3, 0, 1]


def test_rrr_tile_enumeration():
    for ts, vg in (("cccc", [2, 3, 2, 2]), ("aabb", [3, 3, 4, 4]), ("aacc", [4, 4, 2, 3])):
        addrs, grp = oracle.rrr_tile_addresses(ts, vg)
        brute = [a for a in itertools.product(*[range(v) for v in vg])
                 if all(not (a[i] < a[j] and grp[i] == grp[j] and grp[i] != 3)
                        for i in range(4) for j in range(i + 1, 4))]
        assert [tuple(a) for a in addrs] == brute


def test_dltc_task_mul_restatement_vs_einsum():
    # the tile-by-tile restatement of task_mul reproduces the contraction
    O, V, wbO, wbV = 4, 6, 2, 3
    h = {f"h{i}": i - 1 for i in range(1, 13)}
    p = {f"p{i}": 11 + i for i in range(1, 13)}
    # ccsd-t2 ring-like term: R[p3,p4,h1,h2] += T[p3,p5,h1,h6] * I[h6,p4,h2,p5]
    rng = np.random.default_rng(3)
    T = rng.standard_normal((V, V, O, O))
    I = rng.standard_normal((O, V, O, V))
    R = np.zeros((V, V, O, O))
    c_ids = [p["p3"], p["p4"], h["h1"], h["h2"]]
    a_ids = [p["p3"], p["p5"], h["h1"], h["h6"]]
    b_ids = [h["h6"], p["p4"], h["h2"], p["p5"]]
    lo = [0] * 24
    hi = [O // wbO if i < 12 else V // wbV for i in range(24)]
    oracle.dltc_task_mul(c_ids, R, a_ids, T, b_ids, I, 0.5, lo, hi, wbO, wbV)
    ref = 0.5 * np.einsum("aeim,mbje->abij", T, I)
    assert np.allclose(R, ref, rtol=1e-13, atol=1e-13)
    # task_sum: X[p1,p2,h1,h2] += Y[p2,p1,h2,h1]
    Y = rng.standard_normal((V, V, O, O))
    X = np.zeros_like(Y)
    oracle.dltc_task_sum([12, 13, 0, 1], X, [13, 12, 1, 0], Y, 1.0, wbO, wbV)
    assert np.array_equal(X, np.transpose(Y, (1, 0, 3, 2)))


def test_bs_contract_vs_dense_einsum():
    """Block-sparse restatement against the dense einsum of the same tensors
    with the symmetry-forbidden blocks zero: a product of two allowed blocks
    only lands in an allowed C block, so the two agree on every C block."""
    import itertools
    from dtcf_amd import blocksparse as bsp
    sp = bsp.Space([1, 2, 1, 1, 2, 1, 1, 1], [2, 3, 2, 2, 1, 2, 3, 2])
    rng = np.random.default_rng(4)

    def make(kinds):
        t = bsp.BSTensor(sp, kinds, device=None)
        buf = rng.standard_normal(t.size)
        starts = {k: np.concatenate([[0], np.cumsum(sp.counts[k])]) for k in "ov"}
        dense = np.zeros([sum(sp.counts[k]) for k in kinds])
        for g in t.blocks:
            sl = tuple(slice(starts[k][gi], starts[k][gi] + sp.counts[k][gi]) for k, gi in zip(kinds, g))
            off = t.offset[g]
            dense[sl] = buf[off:off + int(np.prod(t.shape[g]))].reshape(t.shape[g])
        meta = {g: (t.offset[g], t.shape[g]) for g in t.blocks}
        return t, buf, dense, meta, starts

    for (ck, ci), (ak, ai), (bk, bi) in ((("oovv", "ijab"), ("oovv", "ikac"), ("oovv", "kjcb")),
                                          (("vvoo", "abij"), ("vvvv", "abcd"), ("vvoo", "cdij"))):
        At, Ab, Ad, Am, _ = make(ak)
        Bt, Bb, Bd, Bm, _ = make(bk)
        Ct, Cb, Cd, Cm, starts = make(ck)
        got = oracle.bs_contract(Cm, Cb.copy(), ci, Am, Ab, ai, Bm, Bb, bi, 0.7)
        dense = Cd + 0.7 * np.einsum(f"{ai},{bi}->{ci}", Ad, Bd)
        for g in Ct.blocks:
            sl = tuple(slice(starts[k][gi], starts[k][gi] + sp.counts[k][gi]) for k, gi in zip(ck, g))
            off = Ct.offset[g]
            blk = got[off:off + int(np.prod(Ct.shape[g]))].reshape(Ct.shape[g])
            np.testing.assert_allclose(blk, dense[sl], rtol=1e-13, atol=1e-13)
        # forbidden C blocks of the dense result stay zero
        for g in itertools.product(range(8), repeat=4):
            if g in Ct.offset:
                continue
            sl = tuple(slice(starts[k][gi], starts[k][gi] + sp.counts[k][gi]) for k, gi in zip(ck, g))
            assert np.abs(dense[sl] - Cd[sl]).max(initial=0.0) < 1e-13


def test_dltc_program_run_vs_loops():
    """A two-layer driver table (an intermediate I built by a multiplication,
    then read by a permuted scaled summation and by a second multiplication)
    against explicit loops: layer order, coefficients, label permutation."""
    rng = np.random.default_rng(9)
    O, V = 2, 3
    table = {
        "tensors": {"T": ["PAR", "PAR", "HOL", "HOL"], "W": ["HOL", "HOL", "HOL", "HOL"],
                    "X": ["HOL", "HOL"], "I_1": ["PAR", "PAR", "HOL", "HOL"],
                    "R": ["PAR", "PAR", "HOL", "HOL"]},
        "muls": [{"c": "I_1", "c_idx": ["p1", "p2", "h1", "h2"], "a": "T", "a_idx": ["p1", "p2", "h3", "h4"],
                  "b": "W", "b_idx": ["h3", "h4", "h1", "h2"], "coef": 0.5},
                 {"c": "R", "c_idx": ["p1", "p2", "h1", "h2"], "a": "I_1", "a_idx": ["p1", "p2", "h3", "h2"],
                  "b": "X", "b_idx": ["h3", "h1"], "coef": -1.0}],
        "sums": [{"c": "R", "c_idx": ["p1", "p2", "h1", "h2"], "a": "I_1", "a_idx": ["p2", "p1", "h2", "h1"],
                  "coef": 2.0}],
        "layers": [[["MUL", 0]], [["SUM", 0], ["MUL", 1]]],
    }
    T = rng.standard_normal((V, V, O, O))
    W = rng.standard_normal((O, O, O, O))
    X = rng.standard_normal((O, O))
    arrays = {"T": T.copy(), "W": W.copy(), "X": X.copy(),
              "I_1": np.zeros((V, V, O, O)), "R": np.zeros((V, V, O, O))}
    oracle.dltc_program_run(table, arrays)
    I = np.zeros((V, V, O, O))
    for a, b, i, j, k, l in itertools.product(range(V), range(V), range(O), range(O), range(O), range(O)):
        I[a, b, i, j] += 0.5 * T[a, b, k, l] * W[k, l, i, j]
    R = np.zeros((V, V, O, O))
    for a, b, i, j in itertools.product(range(V), range(V), range(O), range(O)):
        R[a, b, i, j] += 2.0 * I[b, a, j, i]
        for k in range(O):
            R[a, b, i, j] += -1.0 * I[a, b, k, j] * X[k, i]
    assert np.allclose(arrays["I_1"], I, rtol=0, atol=1e-13)
    assert np.allclose(arrays["R"], R, rtol=0, atol=1e-13)
