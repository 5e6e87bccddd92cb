"""CPU oracle for DTCF's per-rank local contraction path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` import this package, and only as the
checker or as the timed CPU baseline.  The product package ``dtcf_amd`` never
imports it.

``liboracle.so`` (built from ``dtcf_oracle.c`` by ``oracle/Makefile``) holds
the C restatements; this module wraps them with numpy and adds the few pieces
of pure index bookkeeping that are easier to restate in Python:

* ``permutation_maps``   -- contraction.cc:2553-2572 (contr_list) and
  contraction.cc:164-282 (generate_permutation_map).
* ``dltc_index_split``   -- DLTC/expr.cc:13-67 (Multiplication ctor).
* ``rrr_tile_addresses`` -- tensor.cc:652-693 (generate_data order) with
  tensor.cc:617-650 (isValid) and tensor.cc:761-783 (get_bound), one rank.

Each function's pin (what checks it) is named in its docstring.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = {}

_i64 = C.c_int64
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build(ref: bool = False) -> None:
    """Compile liboracle.so (and, when /root/reference exists, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    if ref and os.path.isdir("/root/reference/DLTC"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def lib() -> C.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    L.orc_rrr_transpose.argtypes = [_dp, _dp, _ip, C.c_int, _ip]
    L.orc_rrr_transpose_tiles.argtypes = [_dp, _dp, _ip, C.c_int, _ip, _i64, _i64, C.c_int]
    L.orc_dltc_position.argtypes = [_ip, _ip, C.c_int, _ip]
    L.orc_permute_by_source.argtypes = [_dp, _dp, _ip, C.c_int, _ip]
    L.orc_dltc_transpose.argtypes = [_dp, _ip, _dp, _ip, C.c_int, _ip, C.c_int]
    L.orc_dltc_transpose.restype = C.c_int
    L.orc_block_rank.argtypes = [_ip, C.c_int, _ip]
    L.orc_block_rank.restype = _i64
    L.orc_block_addr.argtypes = [_i64, C.c_int, _ip, _ip]
    L.orc_symm_permutation.argtypes = [_ip, _ip, C.c_int, _ip]
    L.orc_compose_p_maps.argtypes = [_ip, _ip, C.c_int, _ip]
    L.orc_big_matrix_order.argtypes = [C.c_int, _ip, _ip, _ip, _i64, C.c_int, _lp,
                                       C.POINTER(_i64), C.POINTER(_i64)]
    L.orc_big_matrix_order.restype = C.c_int
    L.orc_big_matrix_pack.argtypes = [_dp, _dp, _lp, _i64, _i64, _i64, _i64, _i64, C.c_int]
    L.orc_dgemm.argtypes = [C.c_int, C.c_int, _i64, _i64, _i64, C.c_double, _dp, _i64,
                            _dp, _i64, _dp, _i64]
    L.orc_dgemm_nt_blocked.argtypes = [_i64, _i64, _i64, C.c_double, _dp, _i64, _dp, _i64,
                                       _dp, _i64, C.c_int]
    L.orc_checkM.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int]
    L.orc_checkS.argtypes = [_dp, _dp, C.c_int, C.c_int]
    L.orc_drand48_fill.argtypes = [_dp, _i64, C.POINTER(C.c_uint64)]
    L.orc_denom_update.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int]
    L.orc_energy.argtypes = [_dp, _dp, C.c_int, C.c_int]
    L.orc_energy.restype = C.c_double
    L.orc_norm2.argtypes = [_dp, _i64]
    L.orc_norm2.restype = C.c_double
    _LIB = L
    return L


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


# --------------------------------------------------------------------------
# permutations
# --------------------------------------------------------------------------
def rrr_transpose(tile, pmap, widths):
    """helper.cc:275-522 -- input dim i -> output position pmap[i].
    Pinned by brute force (tests/test_oracle.py)."""
    tile = _f64(tile).ravel()
    out = np.empty_like(tile)
    lib().orc_rrr_transpose(tile, out, _i32(pmap), len(pmap), _i32(widths))
    return out


def rrr_transpose_tiles(tiles, pmap, widths, nthreads=0):
    tiles = _f64(tiles).ravel()
    bs = int(np.prod(widths))
    out = np.empty_like(tiles)
    lib().orc_rrr_transpose_tiles(tiles, out, _i32(pmap), len(pmap), _i32(widths),
                                  tiles.size // bs, bs, nthreads)
    return out


def dltc_position(x, y):
    """DLTC/transp.h:8-17."""
    p = np.zeros(len(x), dtype=np.int32)
    lib().orc_dltc_position(_i32(x), _i32(y), len(x), p)
    return p


def permute_by_source(buf, s, widths):
    buf = _f64(buf).ravel()
    out = np.empty_like(buf)
    lib().orc_permute_by_source(buf, out, _i32(s), len(s), _i32(widths))
    return out


def dltc_transpose(buf_x, x, y, wx, q2_literal=False, out=None):
    """DLTC/transp.cc:3-59.  Returns (status, buf_y).  Pinned against the
    reference's kernels compiled into oracle/_ref."""
    buf_x = _f64(buf_x).ravel()
    buf_y = np.zeros_like(buf_x) if out is None else out
    st = lib().orc_dltc_transpose(buf_x, _i32(x), buf_y, _i32(y), len(x), _i32(wx),
                                  1 if q2_literal else 0)
    return st, buf_y


# --------------------------------------------------------------------------
# block grid / symmetry / big matrix
# --------------------------------------------------------------------------
def block_rank(addr, grid):
    return int(lib().orc_block_rank(_i32(addr), len(grid), _i32(grid)))


def block_addr(rank, grid):
    a = np.zeros(len(grid), dtype=np.int32)
    lib().orc_block_addr(int(rank), len(grid), _i32(grid), a)
    return a


def symm_permutation(addr, sg):
    s = np.zeros(len(addr), dtype=np.int32)
    lib().orc_symm_permutation(_i32(addr), _i32(sg), len(addr), s)
    return s


def compose_p_maps(map1, map2):
    o = np.zeros(len(map1), dtype=np.int32)
    lib().orc_compose_p_maps(_i32(map1), _i32(map2), len(map1), o)
    return o


def big_matrix_order(vgrid, p_map, permuted_addr, nrow_dims):
    """transpose_helper.cc:20-171 tile placement.  Returns (order, nrows, ncols)
    or raises ValueError where the reference asserts (:126)."""
    pa = _i32(permuted_addr).reshape(-1, len(vgrid))
    n = pa.shape[0]
    order = np.zeros(max(n, 1), dtype=np.int64)
    nr, nc = _i64(), _i64()
    st = lib().orc_big_matrix_order(len(vgrid), _i32(vgrid), _i32(p_map), _i32(pa.ravel()),
                                    n, nrow_dims, order, C.byref(nr), C.byref(nc))
    if st != 0:
        raise ValueError("rows of the big matrix hold different column counts")
    return order[: nr.value * nc.value], nr.value, nc.value


def big_matrix_pack(tiles, order, nrows, ncols, rpb, cpb):
    tiles = _f64(tiles).ravel()
    big = np.zeros(nrows * rpb * ncols * cpb, dtype=np.float64)
    lib().orc_big_matrix_pack(tiles, big, _np64(order), nrows, ncols, rpb, cpb, rpb * cpb, 0)
    return big.reshape(nrows * rpb, ncols * cpb)


def big_matrix_unpack(big, order, nrows, ncols, rpb, cpb, ntiles):
    big = _f64(big).ravel()
    tiles = np.zeros(ntiles * rpb * cpb, dtype=np.float64)
    lib().orc_big_matrix_pack(tiles, big, _np64(order), nrows, ncols, rpb, cpb, rpb * cpb, 1)
    return tiles


def _np64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))


# --------------------------------------------------------------------------
# GEMM
# --------------------------------------------------------------------------
def dgemm(at, bt, m, n, k, alpha, A, lda, B, ldb, Cm, ldc):
    """localcontract.cc:3-21 / compute.cc:3-36 intended math, in place on Cm.
    Pinned against numpy float64 matmul."""
    lib().orc_dgemm(int(at), int(bt), m, n, k, float(alpha), _f64(A).ravel(), lda,
                    _f64(B).ravel(), ldb, Cm, ldc)
    return Cm


def dgemm_nt_blocked(m, n, k, alpha, A, lda, B, ldb, Cm, ldc, nthreads=0):
    lib().orc_dgemm_nt_blocked(m, n, k, float(alpha), A, lda, B, ldb, Cm, ldc, nthreads)
    return Cm


# --------------------------------------------------------------------------
# checkers, drand48, CC closed forms
# --------------------------------------------------------------------------
def checkM(A, B, O, V):
    sC = np.zeros(V * V * O * O)
    lib().orc_checkM(_f64(A).ravel(), _f64(B).ravel(), sC, O, V)
    return sC


def checkS(A, O, V):
    sC = np.zeros(V * V * O * O)
    lib().orc_checkS(_f64(A).ravel(), sC, O, V)
    return sC


class Drand48:
    """glibc drand48 restated (default seed 0x1234ABCD330E)."""

    def __init__(self, state=0x1234ABCD330E):
        self.state = C.c_uint64(state)

    def fill(self, n):
        out = np.empty(n, dtype=np.float64)
        lib().orc_drand48_fill(out, n, C.byref(self.state))
        return out


def denom_update(R, eo, ev):
    o, v = len(eo), len(ev)
    T = np.empty(v * v * o * o)
    lib().orc_denom_update(_f64(R).ravel(), T, _f64(eo), _f64(ev), o, v)
    return T


def energy(Vijab, Tabij, o, v):
    return float(lib().orc_energy(_f64(Vijab).ravel(), _f64(Tabij).ravel(), o, v))


def norm2(R):
    R = _f64(R).ravel()
    return float(lib().orc_norm2(R, R.size))


def ccd_iteration(Vvvvv, Vvvoo, W, T, eo, ev):
    """One CCD-like amplitude iteration of the precision study, numpy einsum
    over dense arrays: ladder[a,b,i,j] = 1/2 sum_cd V[a,b,c,d] T[c,d,i,j],
    ring[a,b,i,j] = sum_kc T[a,c,i,k] W[k,j,c,b], R = V_vvoo + ladder + ring,
    T_new = R / D (``denom_update``) and E = 1/4 sum V_oovv[i,j,a,b]
    T_new[a,b,i,j] with V_oovv[i,j,a,b] = V_vvoo[a,b,i,j] (``energy``).
    Returns (ladder, ring, T_new, E), the first three as [v,v,o,o] arrays.
    New (the reference stops at the contraction); pinned by brute-force
    loops on a tiny system (tests/test_oracle.py)."""
    v, o = Vvvoo.shape[0], Vvvoo.shape[2]
    lad = 0.5 * np.einsum("abcd,cdij->abij", Vvvvv, T, optimize=True)
    ring = np.einsum("acik,kjcb->abij", T, W, optimize=True)
    R = Vvvoo + lad + ring
    Tn = denom_update(R, eo, ev).reshape(v, v, o, o)
    E = energy(np.ascontiguousarray(np.transpose(Vvvoo, (2, 3, 0, 1))), Tn, o, v)
    return lad, ring, Tn, E


def dltc_program_run(table, arrays):
    """A DLTC driver's statements (DLTC/ccd-t2.cc, ccsd-t1.cc, ccsd-t2.cc as
    recorded in dtcf_amd/data/dltc_drivers.json) in layer order, numpy:
    summation C += coef * A permuted to C's labels (DLTC/task.cc:51-81),
    multiplication C += coef * sum over the shared labels of A * B
    (task.cc:83-137 with compute.cc:3-36).  ``arrays`` maps every tensor
    name to an array (accumulators at zero) and is updated in place.  Pinned
    by explicit loops over a two-layer table (tests/test_oracle.py) and by
    the brute-force loops of termlist_statement's test (same Einstein
    sums)."""
    for layer in table["layers"]:
        for kind, i in layer:
            if kind == "MUL":
                m = table["muls"][i]
                arrays[m["c"]] += m["coef"] * termlist_statement(
                    "mul", m["c_idx"], [(m["a"], m["a_idx"]), (m["b"], m["b_idx"])], arrays,
                    optimize=True)
            else:
                s_ = table["sums"][i]
                arrays[s_["c"]] += termlist_statement(
                    "sum", s_["c_idx"], [(s_["a"], s_["a_idx"], s_["coef"])], arrays)
    return arrays


# --------------------------------------------------------------------------
# UHF CCSD: amplitude update, correlation energy and residual norms
# --------------------------------------------------------------------------
def uhf_ccsd_step(a, update=True):
    """One Jacobi update of UHF CCSD amplitudes from their residuals and the
    correlation energy of the updated amplitudes.  New (the reference stops
    at the residuals ra_vo / rb_vo, _uhf_t1.eq:143,242, and rab_vvoo /
    raa_vvoo / rbb_vvoo, _uhf_t2.eq:519,769,996).

    ``a`` maps names to float64 arrays: eoa, eob, eva, evb (orbital
    energies), fa_ov [Oa,Va], fb_ov [Ob,Vb], vaa_oovv [Oa,Oa,Va,Va],
    vab_oovv [Oa,Ob,Va,Vb], vbb_oovv [Ob,Ob,Vb,Vb], the amplitudes ta_vo
    [Va,Oa], tb_vo [Vb,Ob], taa_vvoo [Va,Va,Oa,Oa], tab_vvoo [Va,Vb,Oa,Ob],
    tbb_vvoo [Vb,Vb,Ob,Ob] and, with update, the residuals ra_vo, rb_vo,
    raa_vvoo, rab_vvoo, rbb_vvoo of the same shapes.

    Update (R = 0 at convergence; the residuals hold the diagonal Fock terms
    (f_aa - f_ii) t): t_ai += r_ai / (e_i - e_a), t_abij += r_abij /
    (((e_i + e_j) - e_a) - e_b), every spin block with its own orbital
    energies, in that operation order.

    Energy of the (updated) amplitudes, the spin-orbital
    E = sum f_ia t_ai + 1/4 sum <ij||ab> t_abij + 1/2 sum <ij||ab> t_ai t_bj
    resolved into spin blocks:
      E1  = sum fa_ov[i,a] ta[a,i] + sum fb_ov[i,a] tb[a,i]
      Eaa = 1/4 sum vaa_oovv[i,j,a,b] (taa[a,b,i,j] + 2 ta[a,i] ta[b,j])
      Eab =     sum vab_oovv[i,j,a,b] (tab[a,b,i,j] +   ta[a,i] tb[b,j])
      Ebb = 1/4 sum vbb_oovv[i,j,a,b] (tbb[a,b,i,j] + 2 tb[a,i] tb[b,j])
    Returns (new amplitudes dict, out) with out = [E, E1, Eaa, Eab, Ebb,
    |R1|, |R2|, |dT|] (norms 0 without update).  Pinned by the spin-orbital
    embedding (``spin_orbital_energy``), the T = 0 closed form and the fixed
    point of the iteration (tests)."""
    t = {k: np.array(a[k], dtype=np.float64, copy=True)
         for k in ("ta_vo", "tb_vo", "taa_vvoo", "tab_vvoo", "tbb_vvoo")}
    eoa, eob, eva, evb = (np.asarray(a[k], dtype=np.float64) for k in ("eoa", "eob", "eva", "evb"))
    r1 = r2 = d2 = 0.0
    if update:
        for tn, rn, eo, ev in (("ta_vo", "ra_vo", eoa, eva), ("tb_vo", "rb_vo", eob, evb)):
            d = eo[None, :] - ev[:, None]
            dt = a[rn] / d
            t[tn] = t[tn] + dt
            r1 += float(np.sum(a[rn] ** 2))
            d2 += float(np.sum(dt ** 2))
        for tn, rn, e1, e2, v1, v2 in (("taa_vvoo", "raa_vvoo", eoa, eoa, eva, eva),
                                       ("tab_vvoo", "rab_vvoo", eoa, eob, eva, evb),
                                       ("tbb_vvoo", "rbb_vvoo", eob, eob, evb, evb)):
            d = ((e1[None, None, :, None] + e2[None, None, None, :]) - v1[:, None, None, None]) \
                - v2[None, :, None, None]
            dt = a[rn] / d
            t[tn] = t[tn] + dt
            r2 += float(np.sum(a[rn] ** 2))
            d2 += float(np.sum(dt ** 2))
    ta, tb = t["ta_vo"], t["tb_vo"]
    e1 = float(np.einsum("ia,ai->", a["fa_ov"], ta) + np.einsum("ia,ai->", a["fb_ov"], tb))
    eaa = 0.25 * float(np.einsum("ijab,abij->", a["vaa_oovv"], t["taa_vvoo"]
                                 + 2.0 * np.einsum("ai,bj->abij", ta, ta)))
    eab = float(np.einsum("ijab,abij->", a["vab_oovv"], t["tab_vvoo"]
                          + np.einsum("ai,bj->abij", ta, tb)))
    ebb = 0.25 * float(np.einsum("ijab,abij->", a["vbb_oovv"], t["tbb_vvoo"]
                                 + 2.0 * np.einsum("ai,bj->abij", tb, tb)))
    out = np.array([e1 + eaa + eab + ebb, e1, eaa, eab, ebb, np.sqrt(r1), np.sqrt(r2),
                    np.sqrt(d2)])
    return t, out


def spin_orbital_energy(a, t):
    """The CCSD correlation energy in spin orbitals, E = sum f_ia t_ai +
    1/4 sum <ij||ab> t_abij + 1/2 sum <ij||ab> t_ai t_bj, from UHF blocks
    embedded into spin-orbital tensors (occupied alpha then beta, virtual
    alpha then beta): <i(a) j(b)||a(a) b(b)> = vab[i,j,a,b] and its three
    images under the antisymmetry of each pair, amplitudes likewise, the
    same-spin blocks as they are.  Independent of ``uhf_ccsd_step``'s
    block factors, which it pins."""
    Oa, Va = a["fa_ov"].shape
    Ob, Vb = a["fb_ov"].shape
    o, v = Oa + Ob, Va + Vb
    A, B = slice(0, Oa), slice(Oa, o)
    a_, b_ = slice(0, Va), slice(Va, v)
    f = np.zeros((o, v))
    f[A, a_], f[B, b_] = a["fa_ov"], a["fb_ov"]
    t1 = np.zeros((v, o))
    t1[a_, A], t1[b_, B] = t["ta_vo"], t["tb_vo"]
    V = np.zeros((o, o, v, v))
    V[A, A, a_, a_] = a["vaa_oovv"]
    V[B, B, b_, b_] = a["vbb_oovv"]
    vab = a["vab_oovv"]
    V[A, B, a_, b_] = vab
    V[B, A, b_, a_] = np.transpose(vab, (1, 0, 3, 2))
    V[A, B, b_, a_] = -np.transpose(vab, (0, 1, 3, 2))
    V[B, A, a_, b_] = -np.transpose(vab, (1, 0, 2, 3))
    T2 = np.zeros((v, v, o, o))
    T2[a_, a_, A, A] = t["taa_vvoo"]
    T2[b_, b_, B, B] = t["tbb_vvoo"]
    tab = t["tab_vvoo"]
    T2[a_, b_, A, B] = tab
    T2[b_, a_, B, A] = np.transpose(tab, (1, 0, 3, 2))
    T2[a_, b_, B, A] = -np.transpose(tab, (0, 1, 3, 2))
    T2[b_, a_, A, B] = -np.transpose(tab, (1, 0, 2, 3))
    return float(np.einsum("ia,ai->", f, t1) + 0.25 * np.einsum("ijab,abij->", V, T2)
                 + 0.5 * np.einsum("ijab,ai,bj->", V, t1, t1))


# --------------------------------------------------------------------------
# index bookkeeping restated in Python
# --------------------------------------------------------------------------
def permutation_maps(a_idx, b_idx, c_idx, block_range):
    """contraction.cc:2553-2572 + 164-282.

    a_idx/b_idx/c_idx: lists of index names (e.g. ['i','k','a','c']).
    block_range: dict name -> tile extent.  Returns dict with p_map_A/B/C,
    inv_p_map_C, inverse_block_range_C, n_a, n_b, n_k, contr_list.
    Pinned by the closed-form checks in tests (maps are permutations,
    externals keep their relative order, contraction pairs follow contr_list).
    """
    contr = [(i, j) for i, x in enumerate(a_idx) for j, y in enumerate(b_idx) if x == y]
    ca = {i for i, _ in contr}
    cb = {j for _, j in contr}
    p_a = [0] * len(a_idx)
    p_b = [0] * len(b_idx)
    p_c = [0] * len(c_idx)
    inv_c = [0] * len(c_idx)
    inv_br = [0] * len(c_idx)
    n_a = n_b = nk_a = nk_b = 1
    ea = 0
    for i, x in enumerate(a_idx):
        if i not in ca:
            n_a *= block_range[x]
            p_a[i] = ea
            ci = c_idx.index(x)
            p_c[ci] = ea
            inv_br[ea] = block_range[x]
            inv_c[ea] = ci
            ea += 1
        else:
            nk_a *= block_range[x]
    eb = 0
    for j, x in enumerate(b_idx):
        if j not in cb:
            n_b *= block_range[x]
            p_b[j] = eb
            ci = c_idx.index(x)
            p_c[ci] = ea + eb
            inv_br[ea + eb] = block_range[x]
            inv_c[ea + eb] = ci
            eb += 1
        else:
            nk_b *= block_range[x]
    assert nk_a == nk_b
    for t, (i, j) in enumerate(contr):
        p_a[i] = ea + t
        p_b[j] = eb + t
    return dict(p_map_A=p_a, p_map_B=p_b, p_map_C=p_c, inv_p_map_C=inv_c,
                inverse_block_range_C=inv_br, n_a=n_a, n_b=n_b, n_k=nk_a,
                contr_list=contr)


def dltc_index_split(ida, idb, nidx=24):
    """DLTC/expr.cc:13-67: m (A only), n (B only), k (both), each in index
    enum order; at = m+k, bt = k+n, ct = m+n."""
    tab = [0] * nidx
    for x in ida:
        tab[x] = 1
    for x in idb:
        tab[x] += 2
    m = [i for i in range(nidx) if tab[i] == 1]
    n = [i for i in range(nidx) if tab[i] == 2]
    k = [i for i in range(nidx) if tab[i] == 3]
    return dict(m=m, n=n, k=k, at=m + k, bt=k + n, ct=m + n)


def rrr_tile_addresses(tensor_str, vgrid):
    """tensor.cc:652-693 generate_data order for one rank (pgrid all 1):
    odometer over tile addresses, last fastest, bounded per get_bound
    (tensor.cc:761-783), kept iff isValid (tensor.cc:617-650)."""
    dims = len(tensor_str)
    group = []
    for ch in tensor_str:
        group.append(0 if ch in "aA" else 1 if ch in "bB" else 3)
    out = []

    def bound(d, cur):
        ch = tensor_str[d]
        if ch in "cCdD" or d == 0:
            return vgrid[d]
        for i in range(d - 1, -1, -1):
            if (ch in "aA" and tensor_str[i] in "aA") or (ch in "bB" and tensor_str[i] in "bB"):
                return cur[i] + 1
        return vgrid[d]

    def valid(cur):
        for i in range(dims):
            for j in range(i + 1, dims):
                if cur[i] < cur[j] and group[i] == group[j] and group[i] != 3:
                    return False
        return all(cur[i] < vgrid[i] for i in range(dims))

    cur = [0] * dims

    def rec(d):
        for i in range(bound(d, cur)):
            cur[d] = i
            if d == dims - 1:
                if valid(cur):
                    out.append(list(cur))
            else:
                rec(d + 1)

    rec(0)
    return np.array(out, dtype=np.int32).reshape(-1, dims), group


# --------------------------------------------------------------------------
# reference kernels compiled from source (oracle/_ref)
# --------------------------------------------------------------------------
DLTC_4D_NAMES = [
    "ijkl", "ijlk", "ikjl", "iklj", "iljk", "ilkj", "jikl", "jilk", "jkil", "jkli",
    "jlik", "jlki", "kijl", "kilj", "kjil", "kjli", "klij", "klji", "lijk", "likj",
    "ljik", "ljki", "lkij", "lkji",
]


def ref_transp(kind="seq"):
    """The reference's own DLTC transpose kernels (transp-kernel-{seq,omp}.cc)
    compiled by oracle/Makefile into oracle/_ref, or None if not built."""
    if kind in _REF:
        return _REF[kind]
    path = os.path.join(_HERE, "_ref", f"libref_transp_{kind}.so")
    if not os.path.exists(path):
        _REF[kind] = None
        return None
    L = C.CDLL(path)
    fns = {}
    for nm in ["ij", "ji"]:
        f = getattr(L, f"_Z{len('transpose_' + nm)}transpose_{nm}PKdPdii")
        f.argtypes = [_dp, _dp, C.c_int, C.c_int]
        fns[nm] = f
    for nm in DLTC_4D_NAMES:
        f = getattr(L, f"_Z{len('transpose_' + nm)}transpose_{nm}PKdPdiiii")
        f.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int]
        fns[nm] = f
    _REF[kind] = fns
    return fns


# --------------------------------------------------------------------------
# DLTC task_mul / task_sum restated tile by tile (small sizes only)
# --------------------------------------------------------------------------
def dltc_widths(ids, wbO, wbV, nocc=12):
    """getWB (DLTC/define.h:59-65)."""
    return [wbO if x < nocc else wbV for x in ids]


def _box(ids, idvm, wbO, wbV, nocc):
    return tuple(slice(idvm[x] * w, (idvm[x] + 1) * w)
                 for x, w in zip(ids, dltc_widths(ids, wbO, wbV, nocc)))


def _odometer(ids, lo, hi):
    """IndexLoops::next (DLTC/loops.cc:17-31): last index fastest."""
    import itertools
    ranges = [range(lo[x], hi[x]) for x in ids]
    for vals in itertools.product(*ranges):
        yield dict(zip(ids, vals))


def dltc_task_mul(c_ids, Cd, a_ids, Ad, b_ids, Bd, coef, itlb, itub, wbO, wbV, nocc=12,
                  nidx=24):
    """DLTC/task.cc:83-137: for each ext tile of C (odometer over tC.ids),
    buf_ct = 0; for each sum tile: getTile A and B (dltc_tensor.cc:181-267,
    dense boxes), transpose to at_ids / bt_ids (transp.cc, identity copies),
    compute buf_ct += at * bt (compute.cc, row-major m x k by k x n); then
    transpose ct -> C ids and accumulate coef * buf (dltc_tensor.cc:278-305).
    Cd is updated in place (numpy arrays shaped by the index ranges)."""
    sp = dltc_index_split(a_ids, b_ids, nidx)
    at, bt, ct, kk = sp["at"], sp["bt"], sp["ct"], sp["k"]
    m = int(np.prod(dltc_widths(sp["m"], wbO, wbV, nocc))) if sp["m"] else 1
    n = int(np.prod(dltc_widths(sp["n"], wbO, wbV, nocc))) if sp["n"] else 1
    k = int(np.prod(dltc_widths(kk, wbO, wbV, nocc))) if kk else 1
    for ext in _odometer(list(c_ids), itlb, itub):
        buf_ct = np.zeros(m * n)
        for sm in _odometer(kk, itlb, itub):
            idvm = dict(ext)
            idvm.update(sm)
            ta = np.ascontiguousarray(Ad[_box(a_ids, idvm, wbO, wbV, nocc)]).ravel()
            st, at_buf = dltc_transpose(ta, list(a_ids), at, dltc_widths(a_ids, wbO, wbV, nocc))
            assert st == 0
            tb = np.ascontiguousarray(Bd[_box(b_ids, idvm, wbO, wbV, nocc)]).ravel()
            st, bt_buf = dltc_transpose(tb, list(b_ids), bt, dltc_widths(b_ids, wbO, wbV, nocc))
            assert st == 0
            dgemm(0, 0, m, n, k, 1.0, at_buf, k, bt_buf, n, buf_ct, n)
        st, buf_c = dltc_transpose(buf_ct, ct, list(c_ids), dltc_widths(ct, wbO, wbV, nocc))
        assert st == 0
        box = _box(c_ids, ext, wbO, wbV, nocc)
        Cd[box] += coef * buf_c.reshape(Cd[box].shape)
    return Cd


def dltc_task_sum(c_ids, Cd, a_ids, Ad, coef, wbO, wbV, nocc=12):
    """DLTC/task.cc:51-81 over all tiles: tC tile += coef * transpose(tA tile)."""
    lo = {x: 0 for x in c_ids}
    hi = {x: (Cd.shape[i] // dltc_widths([x], wbO, wbV, nocc)[0]) for i, x in enumerate(c_ids)}
    for ext in _odometer(list(c_ids), lo, hi):
        ta = np.ascontiguousarray(Ad[_box(a_ids, ext, wbO, wbV, nocc)]).ravel()
        st, buf = dltc_transpose(ta, list(a_ids), list(c_ids), dltc_widths(a_ids, wbO, wbV, nocc))
        assert st == 0
        box = _box(c_ids, ext, wbO, wbV, nocc)
        Cd[box] += coef * buf.reshape(Cd[box].shape)
    return Cd


# --------------------------------------------------------------------------
# block-sparse (abelian point group) contraction, block by block (numpy)
# --------------------------------------------------------------------------
def bs_contract(c_blocks, Cbuf, c_idx, a_blocks, Abuf, a_idx, b_blocks, Bbuf, b_idx, alpha,
                nirrep=8):
    """C[c_idx] += alpha * A[a_idx] * B[b_idx] over stored blocks.

    *_blocks: dict irrep-tuple -> (offset, shape) of the stored (allowed)
    blocks.  For every C block and every irrep assignment of the contracted
    indices whose A and B blocks are stored, C_block += alpha *
    einsum(A_block, B_block) -- the dense contraction restricted to the
    symmetry-allowed blocks, summed pair by pair.  Pinned in tests by the
    dense einsum of the zero-filled tensors."""
    import itertools
    contr = [x for x in a_idx if x in b_idx]
    spec = f"{a_idx},{b_idx}->{c_idx}"
    for gc, (co, cs) in c_blocks.items():
        irr = dict(zip(c_idx, gc))
        Cb = Cbuf[co:co + int(np.prod(cs))].reshape(cs)
        for gk in itertools.product(range(nirrep), repeat=len(contr)):
            irr.update(zip(contr, gk))
            ga = tuple(irr[x] for x in a_idx)
            gb = tuple(irr[x] for x in b_idx)
            if ga not in a_blocks or gb not in b_blocks:
                continue
            ao, ash = a_blocks[ga]
            bo, bsh = b_blocks[gb]
            Ab = Abuf[ao:ao + int(np.prod(ash))].reshape(ash)
            Bb = Bbuf[bo:bo + int(np.prod(bsh))].reshape(bsh)
            Cb += alpha * np.einsum(spec, Ab, Bb, optimize=True)
    return Cbuf


# --------------------------------------------------------------------------
# term-list statements (the UHF CCSD lists), evaluated with numpy einsum
# --------------------------------------------------------------------------
def _letters(names):
    table = {}
    for n in names:
        if n not in table:
            table[n] = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ"[len(table)]
    return table


def termlist_statement(kind, c_idx, terms, arrays, optimize=False):
    """One statement of a term list.  kind "mul": terms = [(name, idx),
    (name, idx)], result C[c_idx] = sum over the shared indices of A * B (the
    DLTC multiplication, DLTC/expr.cc:13-67 with DLTC/compute.cc:3-36, as a
    plain Einstein sum).  kind "sum": terms = [(name, idx, coef), ...],
    result = sum of coef * X permuted to c_idx (DLTC/task.cc:51-81).  Pinned
    by the brute-force loops of tests/test_termlist.py."""
    if kind == "mul":
        (an, ai), (bn, bi) = terms
        L = _letters(list(c_idx) + list(ai) + list(bi))
        spec = "".join(L[x] for x in ai) + "," + "".join(L[x] for x in bi) + "->" + \
            "".join(L[x] for x in c_idx)
        return np.einsum(spec, arrays[an], arrays[bn], optimize=optimize)
    out = None
    for name, idx, coef in terms:
        perm = [list(idx).index(x) for x in c_idx]
        term = coef * np.transpose(arrays[name], perm)
        out = term if out is None else out + term
    return np.ascontiguousarray(out)


def termlist_run(statements, arrays):
    """Runs statements in order; ``statements`` are (lhs name, kind, c_idx,
    terms) tuples; ``arrays`` maps input names to numpy arrays and receives
    every assigned array."""
    for lhs, kind, c_idx, terms in statements:
        arrays[lhs] = termlist_statement(kind, c_idx, terms, arrays)
    return arrays


def termlist_run_sampled(statements, get_input, shapes, want):
    """The statements of ``termlist_run`` evaluated only on boxes of index
    values: ``want`` maps an output name to one list of index values per
    dimension (None = the whole range).  A backward pass gives every array
    the box its readers need (a contracted dimension needs its whole range,
    an external or summed one the values of the result's box; boxes of one
    array are united per dimension), a forward pass evaluates each statement
    on its box with the same Einstein sums as ``termlist_statement``.
    ``get_input(name)`` returns a whole input array, ``shapes`` maps every
    array name to its extents.  Returns {output name: array over its box}.
    Pinned in tests against ``termlist_run`` on whole small systems."""
    def full(n, d):
        return np.arange(shapes[n][d])

    need = {}

    def add(name, vals):
        cur = need.get(name)
        if cur is None:
            need[name] = [np.unique(v) for v in vals]
        else:
            need[name] = [np.union1d(a, b) for a, b in zip(cur, vals)]

    for name, box in want.items():
        add(name, [full(name, d) if v is None else np.asarray(v) for d, v in enumerate(box)])
    lhs_seen = set()
    for lhs, kind, c_idx, terms in reversed(statements):
        assert lhs not in lhs_seen, f"{lhs} assigned twice"
        lhs_seen.add(lhs)
        if lhs not in need:
            continue
        box = dict(zip(c_idx, need[lhs]))
        for t in terms:
            tname, tidx = t[0], t[1]
            add(tname, [box[x] if x in box else full(tname, d) for d, x in enumerate(tidx)])

    store = {}  # name -> (values per dim, array over the box)

    def take(name, idx, box):
        vals, arr = store[name]
        pos = []
        for d, x in enumerate(idx):
            want_v = box.get(x, vals[d])
            p = np.searchsorted(vals[d], want_v)
            assert np.array_equal(vals[d][p], want_v), (name, d)
            pos.append(p)
        return arr[np.ix_(*pos)]

    for lhs, kind, c_idx, terms in statements:
        for t in terms:
            if t[0] not in store and t[0] in need:
                vals = need[t[0]]
                store[t[0]] = (vals, np.ascontiguousarray(get_input(t[0])[np.ix_(*vals)]))
        if lhs not in need:
            continue
        vals = need[lhs]
        box = dict(zip(c_idx, vals))
        if kind == "mul":
            (an, ai), (bn, bi) = terms
            contr = [x for x in ai if x in bi and x not in c_idx]
            bx = {x: v for x, v in box.items()}
            for x in contr:
                bx.pop(x, None)
            sub = {an: take(an, ai, bx), bn: take(bn, bi, bx)}
            out = termlist_statement("mul", c_idx, [(an, ai), (bn, bi)], sub, optimize=True)
        else:
            sub = {}
            tt = []
            for k, (name, idx, coef) in enumerate(terms):
                key = f"{name}#{k}"
                sub[key] = take(name, idx, box)
                tt.append((key, idx, coef))
            out = termlist_statement("sum", c_idx, tt, sub)
        store[lhs] = (vals, out)
    return {n: store[n][1] for n in want}
