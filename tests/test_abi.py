"""The C-ABI libraries load (no GPU needed) and export every declared symbol."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
PKG = os.path.join(ROOT, "dtcf_amd")


def _declared_c(header):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dtcf_[a-z0-9_]+)\s*\(", text)))


def _exports(lib, demangle=False):
    args = ["nm", "-D", "--defined-only"] + (["-C"] if demangle else []) + [lib]
    out = subprocess.run(args, check=True, capture_output=True, text=True).stdout
    return out


def test_libdtcf_loads_and_exports_every_c_symbol():
    from dtcf_amd import _lib

    L = _lib.lib()
    names = _declared_c("dtcf_local.h")
    assert len(names) >= 20
    ex = _exports(_lib.LIB_PATH)
    for n in names:
        assert re.search(rf"\bT {n}$", ex, re.M), f"{n} declared in dtcf_local.h but not exported"
        getattr(L, n)  # resolvable through ctypes
    assert L.dtcf_abi_version() == 100


def test_dropin_exports_reference_signatures():
    from dtcf_amd import _lib

    lib = _lib.DROPIN_PATH
    assert os.path.exists(lib)
    ex = _exports(lib, demangle=True)
    want = [
        "local_contraction(int, int, int, double*, double*, double*, int, int, double)",
        "transpose(double*&, double*&, int*, int, int*)",
        "transpose_ij(double const*, double*, int, int)",
        "transpose_ji(double const*, double*, int, int)",
        "RRR::Block_Grid::Block_Grid(int, int*)",
        "RRR::Block_Grid::get_block_rank(int*&)",
        "RRR::Block_Grid::get_block_addr(int, int*&)",
        "DLTC::transpose(double*, std::vector<DLTC::IndexName, std::allocator<DLTC::IndexName> > const&, "
        "double*, std::vector<DLTC::IndexName, std::allocator<DLTC::IndexName> > const&)",
    ]
    import oracle

    for nm in oracle.DLTC_4D_NAMES:
        want.append(f"transpose_{nm}(double const*, double*, int, int, int, int)")
    for w in want:
        assert w in ex, f"missing drop-in symbol {w}"
    ex_c = _exports(lib)
    assert re.search(r"\bT dtcf_dltc_set_widths$", ex_c, re.M)


def test_error_path_without_gpu_calls():
    """Argument errors are reported before any device work."""
    import ctypes as C

    from dtcf_amd import _lib

    L = _lib.lib()
    st = L.dtcf_gemm_f64(0, 0, -1, 1, 1, 1.0, None, 1, None, 1, None, 1, None)
    assert st == -1 and b"negative" in L.dtcf_last_error()
    st = L.dtcf_gemm_set_config(17)
    assert st == -1
    d = _lib.Strided()
    d.dims = 9
    st = L.dtcf_permute_copy_f64(None, None, C.byref(d), 1, None, 0, None, 0, 1.0, 0.0, None)
    assert st == -1
    w = _lib.i64_array([2, 3])
    s = _lib.i32_array([0, 0])
    st = L.dtcf_permute_tiles_f64(None, None, 2, w, s, 1, None)
    assert st == -1 and b"permutation" in L.dtcf_last_error()


def test_gemm_group_rejects_bad_problems_without_gpu_calls():
    """dtcf_gemm_group_f64 validates every problem before launching any."""
    import ctypes as C

    from dtcf_amd import _lib, blocksparse as bsp

    L = _lib.lib()
    P = bsp.GemmProblem
    ok = P(4, 4, 4, 16, 4, 16, 4, 16, 4)      # fake, aligned, never dereferenced
    empty = P(4, 0, 4, 16, 4, 16, 4, 16, 4)
    short_pitch = P(4, 4, 8, 16, 4, 16, 8, 16, 4)
    null_b = P(4, 4, 4, 16, 4, None, 4, 16, 4)
    for bad, msg in ((empty, b"empty"), (short_pitch, b"pitch"), (null_b, b"null")):
        arr = (P * 2)(ok, bad)
        st = L.dtcf_gemm_group_f64(C.cast(arr, C.c_void_p), 2, 1.0, 1, 0, None)
        assert st == -1 and msg in L.dtcf_last_error(), msg
    arr = (P * 1)(ok)
    assert L.dtcf_gemm_group_f64(C.cast(arr, C.c_void_p), 1, 1.0, 1, 9, None) == -1
    assert L.dtcf_gemm_group_f64(None, 0, 1.0, 1, 0, None) == 0
    # the layout form checks pitches against its layouts: A k x m needs lda >= m
    ex = L.dtcf_gemm_group_ex_f64
    tall = (P * 1)(P(8, 4, 4, 16, 4, 16, 4, 16, 4))
    assert ex(C.cast(tall, C.c_void_p), 1, 1, 1, 1.0, 1, 0, None) == -1
    assert b"pitch" in L.dtcf_last_error()
    wide_b = (P * 1)(P(4, 8, 4, 16, 4, 16, 4, 16, 8))   # B k x n needs ldb >= n
    assert ex(C.cast(wide_b, C.c_void_p), 1, 0, 0, 1.0, 1, 0, None) == -1
    assert ex(C.cast(arr, C.c_void_p), 1, 2, 1, 1.0, 1, 0, None) == -1
    assert b"transa" in L.dtcf_last_error()


def test_gemm_split_level_one_is_empty_without_gpu_calls():
    """max_level 1 keeps no slice pair (i + j >= 2): C is left alone and no
    work is enqueued (the flush interval would divide by a zero level
    bound); bad precisions are refused first."""
    from dtcf_amd import mixed
    from dtcf_amd._lib import lib

    L = lib()
    fake = 256  # aligned, never dereferenced
    for prec, ns in ((mixed.I8, 4), (mixed.BF16X3, 3), (mixed.F16X3, 3)):
        st = L.dtcf_gemm_split_f64(prec, ns, 64, 64, 64, 1.0, fake, fake, fake, fake, 1, fake,
                                   None, 64, None)
        assert st == 0, L.dtcf_last_error()
    st = L.dtcf_gemm_split_f64(9, 4, 64, 64, 64, 1.0, fake, fake, fake, fake, 0, fake, None, 64,
                               None)
    assert st == -1 and b"precision" in L.dtcf_last_error()
