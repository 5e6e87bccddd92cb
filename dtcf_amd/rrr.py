"""RRR tensors held as tiles, and one rank's local contraction step.

``TileTensor`` mirrors the tile storage of ``RRR::Tensor`` (tensor.h:29-421)
for one rank: the tile list in generate_data order (tensor.cc:652-693,
odometer over tile addresses, symmetric dims bounded by get_bound
tensor.cc:761-783, kept iff isValid tensor.cc:617-650), each tile row-major
with extents ``block_range = size / vgrid`` (tensor.cc:27-32), all tiles
resident in one device buffer.

``LocalStep`` wraps ``dtcf_step_plan/run`` (the C++ planner in
csrc/local_step.hip), the per-rank work of ``Contraction::contract`` around
``transpose_and_dgemm`` (contraction.cc:1779-1976, 2692-2830).
"""
from __future__ import annotations

import ctypes as C
import itertools

import numpy as np
import torch

from . import _lib
from ._lib import check, lib

LITERAL, PACKED, FUSED = 0, 1, 2


class Tileset(C.Structure):
    _fields_ = [
        ("dims", C.c_int32),
        ("vgrid", C.c_int32 * _lib.DTCF_MAX_DIMS),
        ("block_range", C.c_int32 * _lib.DTCF_MAX_DIMS),
        ("sym_group", C.c_int32 * _lib.DTCF_MAX_DIMS),
        ("ntiles", C.c_int64),
        ("tile_addr", C.POINTER(C.c_int32)),
        ("tiles", C.c_void_p),
    ]


def sym_groups(tensor_str: str):
    """tensor.cc:93-164: 'a'/'A' -> group 0, 'b'/'B' -> 1, others non-symmetric (3)."""
    return [0 if ch in "aA" else 1 if ch in "bB" else 3 for ch in tensor_str]


def tile_addresses(tensor_str: str, vgrid):
    """Tile addresses of one rank (pgrid all 1) in generate_data order."""
    dims = len(tensor_str)
    grp = sym_groups(tensor_str)
    out = []
    for addr in itertools.product(*[range(v) for v in vgrid]):
        ok = True
        for i in range(dims):
            for j in range(i + 1, dims):
                if addr[i] < addr[j] and grp[i] == grp[j] and grp[i] != 3:
                    ok = False
        if ok:
            out.append(addr)
    return np.array(out, dtype=np.int32).reshape(-1, dims)


class TileTensor:
    """Tiles of one rank's share of an RRR tensor, resident on the device."""

    def __init__(self, tensor_str, size, vgrid, device="cuda", data=None, addrs=None):
        self.tensor_str = tensor_str
        self.dims = len(tensor_str)
        self.size = list(size)
        self.vgrid = list(vgrid)
        assert all(s % v == 0 for s, v in zip(size, vgrid)), "size must be a multiple of vgrid"
        self.block_range = [s // v for s, v in zip(size, vgrid)]
        self.block_size = int(np.prod(self.block_range))
        self.addrs = tile_addresses(tensor_str, vgrid) if addrs is None else np.asarray(addrs, np.int32)
        self.ntiles = len(self.addrs)
        if data is None:
            self.tiles = torch.zeros(self.ntiles * self.block_size, dtype=torch.float64, device=device)
        else:
            self.tiles = data
        assert self.tiles.numel() == self.ntiles * self.block_size

    @staticmethod
    def from_dense(tensor_str, dense: np.ndarray, vgrid, device="cuda"):
        """Cut a dense array into tiles (host), in generate_data order."""
        t = TileTensor(tensor_str, dense.shape, vgrid, device="cpu")
        buf = tiles_from_dense(dense, vgrid, t.addrs)
        t.tiles = torch.from_numpy(buf).to(device)
        return t

    def tileset(self) -> Tileset:
        ts = Tileset()
        ts.dims = self.dims
        for d in range(self.dims):
            ts.vgrid[d] = self.vgrid[d]
            ts.block_range[d] = self.block_range[d]
        for d, g in enumerate(sym_groups(self.tensor_str)):
            ts.sym_group[d] = g
        ts.ntiles = self.ntiles
        self._addr_c = np.ascontiguousarray(self.addrs.ravel(), dtype=np.int32)
        ts.tile_addr = self._addr_c.ctypes.data_as(C.POINTER(C.c_int32))
        ts.tiles = self.tiles.data_ptr() if self.ntiles else None
        return ts


def tiles_from_dense(dense: np.ndarray, vgrid, addrs) -> np.ndarray:
    br = [s // v for s, v in zip(dense.shape, vgrid)]
    bs = int(np.prod(br))
    out = np.empty(len(addrs) * bs, dtype=np.float64)
    for t, a in enumerate(addrs):
        sl = tuple(slice(a[d] * br[d], (a[d] + 1) * br[d]) for d in range(len(br)))
        out[t * bs:(t + 1) * bs] = dense[sl].ravel()
    return out


def dense_from_tiles(tiles: np.ndarray, shape, vgrid, addrs) -> np.ndarray:
    br = [s // v for s, v in zip(shape, vgrid)]
    bs = int(np.prod(br))
    out = np.zeros(shape, dtype=np.float64)
    for t, a in enumerate(addrs):
        sl = tuple(slice(a[d] * br[d], (a[d] + 1) * br[d]) for d in range(len(br)))
        out[sl] = tiles[t * bs:(t + 1) * bs].reshape(br)
    return out


class LocalStep:
    """One rank's local step C += alpha * A * B (index strings as
    Contraction::contract takes them, e.g. "i,k,a,c")."""

    def __init__(self, A: TileTensor, a_idx, B: TileTensor, b_idx, C_: TileTensor, c_idx,
                 mode=FUSED):
        self._keep = (A, B, C_)
        ta, tb, tc = A.tileset(), B.tileset(), C_.tileset()
        self._ts = (ta, tb, tc)
        h = C.c_void_p()
        check(lib().dtcf_step_plan(C.byref(ta), a_idx.encode(), C.byref(tb), b_idx.encode(),
                                   C.byref(tc), c_idx.encode(), int(mode), C.byref(h)))
        self.h = h
        t = (C.c_int64 * 3)()
        b = (C.c_int64 * 3)()
        check(lib().dtcf_step_info(h, t, b))
        self.tile_mnk = tuple(t)
        self.big_mnk = tuple(b)

    @property
    def flops(self):
        M, N, K = self.big_mnk
        return 2.0 * M * N * K

    def run(self, alpha=1.0):
        check(lib().dtcf_step_run(self.h, float(alpha),
                                  C.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def enable_timing(self, max_runs: int):
        check(lib().dtcf_step_enable_timing(self.h, int(max_runs)))

    def kernel_times_ms(self, max_n: int = 4096):
        buf = (C.c_float * max_n)()
        n = C.c_int()
        check(lib().dtcf_step_kernel_times(self.h, buf, max_n, C.byref(n)))
        return [float(buf[i]) for i in range(n.value)]

    def __del__(self):
        try:
            if self.h:
                lib().dtcf_step_destroy(self.h)
                self.h = None
        except Exception:
            pass


def _declare_step():
    L = lib()
    _vp = C.c_void_p
    L.dtcf_step_plan.argtypes = [_vp, C.c_char_p, _vp, C.c_char_p, _vp, C.c_char_p, C.c_int,
                                 C.POINTER(_vp)]
    L.dtcf_step_plan.restype = C.c_int
    L.dtcf_step_run.argtypes = [_vp, C.c_double, _vp]
    L.dtcf_step_run.restype = C.c_int
    L.dtcf_step_info.argtypes = [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.dtcf_step_info.restype = C.c_int
    L.dtcf_step_destroy.argtypes = [_vp]
    L.dtcf_step_destroy.restype = C.c_int
    L.dtcf_step_enable_timing.argtypes = [_vp, C.c_int]
    L.dtcf_step_enable_timing.restype = C.c_int
    L.dtcf_step_kernel_times.argtypes = [_vp, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int)]
    L.dtcf_step_kernel_times.restype = C.c_int


try:
    _declare_step()
except ImportError:
    pass
