"""The reference's hard-coded DLTC CC drivers as programs on the GPU.

DLTC/ccd-t2.cc, DLTC/ccsd-t1.cc and DLTC/ccsd-t2.cc each declare a list of
summations (C += coef * A) and multiplications (C += coef * A * B) over
named tensors with relabelled indices, and run them in layers separated by
barriers: every statement of a layer accumulates into tensors that start at
zero (intermediates I_*, the residual R), and the statements of a layer only
read what earlier layers completed.  dtcf_amd/data/dltc_drivers.json holds
those equations (written by dtcf_amd/data/extract_dltc_drivers.py); a
``Driver`` allocates every tensor dense on the device (occupied extent O,
virtual V), fills the inputs (F_*, V_*, T_*) with seeded normals and runs
the layers, each statement as one whole-term GEMM (``dltc.Term.run``) or as
the reference's finest tasks, one (output tile, summed tile) box each
(``tasks=True``, DLTC/exec.cc:58-110 granularity 6).
"""
from __future__ import annotations

import itertools
import json
import os
import zlib

import torch

from . import dltc, mixed

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "dltc_drivers.json")
DRIVERS = ("ccd_t2", "ccsd_t1", "ccsd_t2")


def load(name):
    with open(_DATA) as f:
        return json.load(f)[name]


def is_accumulator(name):
    """Tensors a driver assigns: the residual and the intermediates."""
    return name == "R" or name.startswith("R_") or name.startswith("I_")


class Driver:
    def __init__(self, name, O, V, wbO, wbV, seed=1, sigma=0.05, device="cuda"):
        self.name = name
        self.table = load(name)
        self.O, self.V, self.wbO, self.wbV = O, V, wbO, wbV
        self.space = dltc.make_space(O, V, wbO, wbV)
        self.shape = {X: tuple(O if k == "HOL" else V for k in kinds)
                      for X, kinds in self.table["tensors"].items()}
        self.t = {}
        for X, shp in self.shape.items():
            n = 1
            for e in shp:
                n *= e
            if is_accumulator(X):
                self.t[X] = torch.zeros(n, dtype=torch.float64, device=device)
            else:
                s = (seed * 1000003 + zlib.crc32(X.encode())) & 0xFFFFFFFF
                self.t[X] = mixed.fill_normal(n, s, sigma, device)
        self.muls = [dltc.Term(self.space, m["c_idx"], self.t[m["c"]], m["a_idx"], self.t[m["a"]],
                               m["b_idx"], self.t[m["b"]], m["coef"]) for m in self.table["muls"]]
        torch.cuda.synchronize()

    def reset(self):
        for X, t in self.t.items():
            if is_accumulator(X):
                t.zero_()

    def _ntiles(self, idx):
        return [self.O // self.wbO if i.startswith("h") else self.V // self.wbV for i in idx]

    def _boxes(self, ext, summed):
        """Tile boxes of one (output tile, summed tile) each, as itlb / itub."""
        for e in itertools.product(*[range(n) for n in self._ntiles(ext)]):
            for s in itertools.product(*[range(n) for n in self._ntiles(summed)]):
                lb, ub = [0] * dltc.NIDX, [0] * dltc.NIDX
                for x, v in zip(list(ext) + list(summed), e + s):
                    lb[dltc.IDX[x]], ub[dltc.IDX[x]] = v, v + 1
                yield lb, ub

    def run_statement(self, kind, i, tasks=False):
        if kind == "MUL":
            m, term = self.table["muls"][i], self.muls[i]
            if not tasks:
                term.run()
                return
            summed = [x for x in m["a_idx"] if x in m["b_idx"] and x not in m["c_idx"]]
            for lb, ub in self._boxes(m["c_idx"], summed):
                term.task(lb, ub)
            return
        s = self.table["sums"][i]
        if not tasks:
            dltc.task_sum(self.space, s["c_idx"], self.t[s["c"]], s["a_idx"], self.t[s["a"]], s["coef"])
            return
        for lb, ub in self._boxes(s["c_idx"], []):
            dltc.task_sum(self.space, s["c_idx"], self.t[s["c"]], s["a_idx"], self.t[s["a"]],
                          s["coef"], lb, ub)

    def run(self, tasks=False):
        """All layers in order (the barriers of the reference become stream
        order: one stream runs every statement)."""
        for layer in self.table["layers"]:
            for kind, i in layer:
                self.run_statement(kind, i, tasks)

    def flops(self):
        return sum(t.flops for t in self.muls)

    def host(self, X):
        return self.t[X].cpu().numpy().reshape(self.shape[X])
