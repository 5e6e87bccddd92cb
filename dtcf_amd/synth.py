"""Seeded synthetic inputs for the local-contraction workloads.

Counter-based generator: element e of a tensor with seed s gets
``splitmix64(s * 2^40 + pair(e))`` style uniforms turned into a normal by
Box-Muller (Box & Muller 1958), so any slice can be generated independently
and the same bytes come out on every host.  The generator replaces the
reference's constant-per-tile values ``sum_d sin(tile_addr[d])``
(tensor.cc:744-758), which hide intra-tile permutation errors, and its
``2*drand48()-1`` checker inputs (DLTC/checkerM.cc:25-26).

Antisymmetrisation under a<->b and i<->j keeps sigma:
``X <- (X - X^{ab} - X^{ij} + X^{ab,ij}) / 2``.
"""
from __future__ import annotations

import numpy as np

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_GOLD = np.uint64(0x9E3779B97F4A7C15)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser (Steele, Lea & Flood, OOPSLA 2014)."""
    with np.errstate(over="ignore"):
        z = x + _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def normal(n: int, seed: int, sigma: float = 1.0, start: int = 0, chunk: int = 1 << 24) -> np.ndarray:
    """n normals N(0, sigma^2) for element indices [start, start+n)."""
    out = np.empty(n, dtype=np.float64)
    base = np.uint64(seed) << np.uint64(40)
    for c0 in range(0, n, chunk):
        e = np.arange(start + c0, start + min(n, c0 + chunk), dtype=np.uint64)
        pair = e >> np.uint64(1)
        h1 = _splitmix64(base + (pair << np.uint64(1)))
        h2 = _splitmix64(base + (pair << np.uint64(1)) + np.uint64(1))
        u1 = ((h1 >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
        u2 = (h2 >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        r = np.sqrt(-2.0 * np.log(u1))
        t = 2.0 * np.pi * u2
        odd = (e & np.uint64(1)).astype(bool)
        out[c0:c0 + len(e)] = sigma * np.where(odd, r * np.sin(t), r * np.cos(t))
    return out


def uniform(n: int, seed: int, lo: float, hi: float) -> np.ndarray:
    base = np.uint64(seed) << np.uint64(40)
    e = np.arange(n, dtype=np.uint64)
    h = _splitmix64(base + e + np.uint64(1 << 39))
    u = (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def antisymmetrize(x: np.ndarray, pairs=((0, 1), (2, 3))) -> np.ndarray:
    """(X - X^{p} - X^{q} + X^{pq}) / 2 for the two index pairs of a 4-D tensor."""
    (a, b), (c, d) = pairs
    pa = list(range(4))
    pa[a], pa[b] = pa[b], pa[a]
    pc = list(range(4))
    pc[c], pc[d] = pc[d], pc[c]
    xa = np.transpose(x, pa)
    xc = np.transpose(x, pc)
    xac = np.transpose(xa, pc)
    return 0.5 * (x - xa - xc + xac)


def fock_diagonals(o: int, v: int, seed: int = 7):
    """Occupied eps ~ U[-2, -0.5], virtual eps ~ U[0.2, 3]."""
    return uniform(o, seed, -2.0, -0.5), uniform(v, seed + 1, 0.2, 3.0)
