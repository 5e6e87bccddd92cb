"""UHF CCSD iterations on one GPU: the T1 and T2 residual programs of the
UHF term lists (``_uhf_t1.eq``, ``_uhf_t2.eq``) served from device-resident
tensors, then the FP64 amplitude update, the correlation energy and the
residual norms (``dtcf_uhf_ccsd_step_f64``).  The reference stops at the
residuals (ra_vo / rb_vo, _uhf_t1.eq:143,242; rab_vvoo / raa_vvoo /
rbb_vvoo, _uhf_t2.eq:519,769,996); the loop around them replaces the
layered per-iteration drivers of DLTC (DLTC/ccsd-t2.cc, ccd-t2.cc:97-131).

One iteration: both residual programs are submitted as requests (they run
concurrently on the server's streams), then
    t_ai   += r_ai   / (e_i - e_a)
    t_abij += r_abij / (((e_i + e_j) - e_a) - e_b)
per spin block, and E(T_new).  The residual programs may run a mixed
precision (digit slices of INT8 MFMA products, recombined in FP64); the
update, the energy and the norms are FP64.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np
import torch

from . import mixed, serve, termlist
from ._lib import check, lib

AMPS = ("ta_vo", "tb_vo", "taa_vvoo", "tab_vvoo", "tbb_vvoo")
RESID = {"ta_vo": "ra_vo", "tb_vo": "rb_vo", "taa_vvoo": "raa_vvoo", "tab_vvoo": "rab_vvoo",
         "tbb_vvoo": "rbb_vvoo"}
OUT_NAMES = ("E", "E1", "Eaa", "Eab", "Ebb", "r1", "r2", "dt")


class UHFDesc(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("Oa", "Ob", "Va", "Vb")] + \
        [(n, C.c_void_p) for n in ("eoa", "eob", "eva", "evb", "fa_ov", "fb_ov", "vaa_oovv",
                                   "vab_oovv", "vbb_oovv", "ra", "rb", "raa", "rab", "rbb",
                                   "ta", "tb", "taa", "tab", "tbb")]


def _declare():
    L = lib()
    L.dtcf_uhf_ccsd_step_f64.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    L.dtcf_uhf_ccsd_step_f64.restype = C.c_int


_declare()


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def step(desc: UHFDesc, update: bool, out: torch.Tensor):
    """Enqueues the update (or, update False, only the energy) on torch's
    current stream; ``out`` (8 float64 on the device) receives
    E, E1, Eaa, Eab, Ebb, |R1|, |R2|, |dT|."""
    check(lib().dtcf_uhf_ccsd_step_f64(C.byref(desc), 1 if update else 0,
                                       C.c_void_p(out.data_ptr()), _stream()))
    return out


def make_desc(sizes, ptr, res=None):
    """Descriptor over device pointers: ``ptr`` maps eoa, eob, eva, evb,
    fa_ov, fb_ov, vaa_oovv, vab_oovv, vbb_oovv and the amplitudes to
    addresses, ``res`` the residual names (ra_vo, ...) to addresses."""
    d = UHFDesc()
    d.Oa, d.Ob, d.Va, d.Vb = (int(sizes[k]) for k in ("Oa", "Ob", "Va", "Vb"))
    for k in ("eoa", "eob", "eva", "evb", "fa_ov", "fb_ov", "vaa_oovv", "vab_oovv", "vbb_oovv"):
        setattr(d, k, ptr[k])
    for amp, short in zip(AMPS, ("ta", "tb", "taa", "tab", "tbb")):
        setattr(d, short, ptr[amp])
    if res is not None:
        for amp, short in zip(AMPS, ("ra", "rb", "raa", "rab", "rbb")):
            setattr(d, short, res[RESID[amp]])
    return d


class CCSystem(serve.System):
    """A synthetic UHF system for CC iterations, resident in a server arena.

    Orbital energies as the precision study draws them: occupied
    U[-2, -0.5], virtual U[0.2, 3], separately per spin.  The Fock blocks
    f_oo, f_vv hold them on the diagonal plus a symmetric off-diagonal part
    N(0, fock_sigma^2); f_ov is N(0, fock_sigma^2) and f_vo its transpose.
    Integrals are N(0, sigma^2), antisymmetrised in every index pair their
    declaration names (each pair (X - X^pq) / sqrt 2, which keeps sigma),
    with v_vvoo[a,b,i,j] = v_oovv[i,j,a,b] (real orbitals).  Amplitudes start
    at zero, so the first iteration gives t_ai = f_ai / D_ai and
    t_abij = v_abij / D_abij."""

    def __init__(self, server, o, v, seed=7, sigma=0.05, fock_sigma=0.01, sizes=None, lists=None):
        sizes = sizes or serve.system_sizes(o, v)
        self.sigma, self.fock_sigma = sigma, fock_sigma
        self.eps = {
            "eoa": mixed.fill_uniform(sizes["Oa"], seed * 8 + 1, -2.0, -0.5),
            "eob": mixed.fill_uniform(sizes["Ob"], seed * 8 + 2, -2.0, -0.5),
            "eva": mixed.fill_uniform(sizes["Va"], seed * 8 + 3, 0.2, 3.0),
            "evb": mixed.fill_uniform(sizes["Vb"], seed * 8 + 4, 0.2, 3.0),
        }
        lists = lists or [termlist.load("t1"), termlist.load("t2")]
        super().__init__(server, o, v, lists, seed=seed, sizes=sizes)

    def _normal(self, name, shp, sigma):
        return mixed.fill_normal(int(np.prod(shp)), self.name_seed(name), sigma).view(*shp)

    def make(self, name, decl, shp):
        spin = name[1]
        if name.startswith("t"):
            return torch.zeros(int(np.prod(shp)), dtype=torch.float64, device="cuda")
        if name.startswith("f"):
            block = name.split("_")[1]
            if block in ("oo", "vv"):
                eps = self.eps[("eo" if block == "oo" else "ev") + spin]
                N = self._normal(name, shp, 1.0)
                N = (N + N.t()) / math.sqrt(2.0)
                N.fill_diagonal_(0.0)
                return (torch.diag(eps) + self.fock_sigma * N).reshape(-1).contiguous()
            if block == "ov":
                return self._normal(name, shp, self.fock_sigma).reshape(-1)
            # f_vo = f_ov^T
            ov = self._normal(f"f{spin}_ov", (shp[1], shp[0]), self.fock_sigma)
            return ov.t().contiguous().reshape(-1)
        if name.endswith("_vvoo"):
            src = name[:-5] + "_oovv"  # v_vvoo[a,b,i,j] = v_oovv[i,j,a,b]
            oovv = self.make(src, self.decl.get(src) or _oovv_decl(name), (shp[2], shp[3], shp[0], shp[1]))
            return oovv.view(shp[2], shp[3], shp[0], shp[1]).permute(2, 3, 0, 1).contiguous().reshape(-1)
        X = self._normal(name, shp, self.sigma)
        for g in decl.groups:
            p, q = g
            assert shp[p] == shp[q]
            X = (X - X.transpose(p, q)) / math.sqrt(2.0)
        return X.contiguous().reshape(-1)


def _oovv_decl(vvoo_name):
    """Declaration of v*_oovv when only v*_vvoo is an input of the lists:
    same-spin blocks are antisymmetric in both pairs."""
    same = vvoo_name[1] == vvoo_name[2]
    return termlist.Array(vvoo_name[:-5] + "_oovv", (), 2, ((0, 1), (2, 3)) if same else ())


class UHFCCSD:
    """CCSD iterations over a ``CCSystem``.  prec 0: FP64 residual programs;
    otherwise a mixed precision for the contractions that take it
    (``serve.Program``), the integral inputs split once and kept resident
    when ``resident_integrals``."""

    def __init__(self, server: serve.Server, system: CCSystem, prec=0, nslice=5, max_level=0,
                 mixed_min=None, resident_integrals=True):
        self.server, self.system = server, system
        self.t1, self.t2 = termlist.load("t1"), termlist.load("t2")
        self.progs = []
        for tl in (self.t1, self.t2):
            res = serve.integral_inputs(tl) if (prec and resident_integrals) else ()
            self.progs.append(serve.Program(server, tl, system, prec=prec, nslice=nslice,
                                            max_level=max_level, mixed_min=mixed_min,
                                            resident=res))
        self.res = {}
        self.outs = []
        for p in self.progs:
            bufs = p.new_outputs()
            self.outs.append(bufs)
            self.res.update(dict(zip(p.outputs, bufs)))
        ptr = dict(system.ptr)
        ptr.update({k: t.data_ptr() for k, t in system.eps.items()})
        self.desc = make_desc(system.sizes, ptr, {k: t.data_ptr() for k, t in self.res.items()})
        self.out = torch.zeros(8, dtype=torch.float64, device="cuda")
        self.flops = sum(p.flops for p in self.progs)

    def residuals(self):
        """Evaluates R1 and R2 at the current amplitudes (both programs at
        once); returns the wall time in seconds."""
        torch.cuda.synchronize()
        t = time.perf_counter()
        rids = [p.submit(outs)[0] for p, outs in zip(self.progs, self.outs)]
        for rid in rids:
            check(self.server.wait(rid, 600000.0))
        dt = time.perf_counter() - t
        for rid in rids:
            self.server.release(rid)
        return dt

    def step(self, update=True):
        """Update (or only evaluate the energy) and return the outputs."""
        step(self.desc, update, self.out)
        vals = self.out.cpu().numpy()  # synchronises: the next requests read the new amplitudes
        return dict(zip(OUT_NAMES, (float(x) for x in vals)))

    def iterate(self):
        t = time.perf_counter()
        tr = self.residuals()
        rec = self.step(True)
        rec["seconds"] = time.perf_counter() - t
        rec["seconds_residuals"] = tr
        rec["tflops_residuals"] = self.flops / tr / 1e12
        return rec

    def energy(self):
        return self.step(False)

    def amplitudes(self):
        return {n: self.system.host(n) for n in AMPS}

    def residual_arrays(self):
        return {n: t.cpu().numpy().reshape(s) for p, outs in zip(self.progs, self.outs)
                for n, t, s in zip(p.outputs, outs, p.out_shapes)}

    def destroy(self):
        for p in self.progs:
            p.destroy()
        self.progs = []


MODES = {
    "fp64": (0, 0, 0),
    "int8x5": (mixed.I8, 5, 0),
    "int8x4/L6": (mixed.I8, 4, 6),
    "int8x4/L5": (mixed.I8, 4, 5),
    "int8x4/L4": (mixed.I8, 4, 4),
    "int8x3/L4": (mixed.I8, 3, 4),
}


def study(o=64, v=256, iters=3, modes=("fp64", "int8x5", "int8x4/L6"), budget_gb=240.0, nslots=1,
          seed=11, sigma=0.05, log=print):
    """Full UHF CCSD iterations, FP64 against mixed residual programs.

    Starting from the first-order amplitudes (one FP64 iteration from zero),
    each of ``iters`` iterations runs every mixed mode from the FP64
    trajectory's amplitudes (restored before each run), then the FP64
    iteration itself, so every mode's error is that of one iteration from the
    same state.  Per mode: seconds per iteration (both residual programs, then
    the update / energy step), contraction TFLOP/s (dense flops of the two
    programs over the residual time), speedup over FP64, and the largest
    |dE| (total and per component E1, Eaa, Eab, Ebb) and max |dR| / max |R|
    per residual block over the iterations.  One server slot by default: the
    programs' large kernels fill the GPU by themselves, and at o=96 v=384 one
    slot ran both forms fastest (FP64 1.242 s, INT8x3/L4 0.320 s per
    iteration, against 1.287 s and 0.328 s on eight)."""
    srv = serve.Server(int(budget_gb * (1 << 30)), nslots)
    try:
        t = time.perf_counter()
        system = CCSystem(srv, o, v, seed=seed, sigma=sigma)
        torch.cuda.empty_cache()
        log(f"system o={o} v={v}: inputs {system.bytes() / 1e9:.2f} GB in "
            f"{time.perf_counter() - t:.1f} s")
        drv = {}
        for m in modes:
            prec, ns, lev = MODES[m]
            drv[m] = UHFCCSD(srv, system, prec=prec, nslice=ns, max_level=lev)
            log(f"{m}: programs planned; resident {srv.usage()['resident'] / 1e9:.2f} GB")
        ref = drv["fp64"]
        first = ref.iterate()
        log(f"first-order amplitudes: E {first['E']:.10f}")
        rec = {m: {"seconds": [], "seconds_residuals": [], "E": [], "dE": [], "dR": []} for m in modes}
        for it in range(iters):
            start = {n: torch.empty(int(np.prod(system.shape[n])), dtype=torch.float64,
                                    device="cuda") for n in AMPS}
            for n, x in start.items():
                serve.copy_device(x.data_ptr(), system.ptr[n], x.numel())
            results, resid = {}, {}
            for m in [x for x in modes if x != "fp64"] + ["fp64"]:
                for n, x in start.items():
                    serve.copy_device(system.ptr[n], x.data_ptr(), x.numel())
                torch.cuda.synchronize()
                r = drv[m].iterate()
                results[m] = r
                resid[m] = drv[m].res
                rec[m]["seconds"].append(r["seconds"])
                rec[m]["seconds_residuals"].append(r["seconds_residuals"])
                rec[m]["E"].append(r["E"])
            e64 = results["fp64"]
            for m in modes:
                dE = {k: abs(results[m][k] - e64[k]) for k in ("E", "E1", "Eaa", "Eab", "Ebb")}
                dR = {n: ((resid[m][n] - resid["fp64"][n]).abs().max() /
                          resid["fp64"][n].abs().max()).item() for n in resid[m]}
                rec[m]["dE"].append(dE)
                rec[m]["dR"].append(dR)
            log(f"iteration {it}: E(fp64) {e64['E']:.10f}, |R2| {e64['r2']:.3e}; " + ", ".join(
                f"{m} {results[m]['seconds']:.3f} s dE {rec[m]['dE'][-1]['E']:.1e}" for m in modes))
            del start
        out = {"o": o, "v": v, "sigma": sigma, "iterations": iters, "flops_per_iteration": ref.flops,
               "energy_first_order": first["E"], "modes": {}}
        t64 = float(np.median(rec["fp64"]["seconds"]))
        for m in modes:
            r = rec[m]
            s = float(np.median(r["seconds"]))
            sr = float(np.median(r["seconds_residuals"]))
            out["modes"][m] = {
                "seconds_per_iteration": s, "seconds_residuals": sr,
                "tflops_residuals": ref.flops / sr / 1e12, "speedup": t64 / s,
                "energy": r["E"],
                "max_energy_err": {k: max(d[k] for d in r["dE"]) for k in r["dE"][0]},
                "max_residual_rel_err": {k: max(d[k] for d in r["dR"]) for k in r["dR"][0]},
            }
        out["usage_gb"] = {k: v / 1e9 for k, v in srv.usage().items()}
        for d in drv.values():
            d.destroy()
        return out
    finally:
        srv.close()


def main():
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser()
    ap.add_argument("--o", type=int, default=64)
    ap.add_argument("--v", type=int, default=256)
    ap.add_argument("--iters", type=int, default=3)
    ap.add_argument("--budget-gb", type=float, default=240.0)
    ap.add_argument("--sigma", type=float, default=0.05)
    ap.add_argument("--modes", default="fp64,int8x5,int8x4/L6")
    ap.add_argument("--nslots", type=int, default=1)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    check(lib().dtcf_init(0))
    res = study(args.o, args.v, args.iters, tuple(args.modes.split(",")), args.budget_gb,
                nslots=args.nslots, sigma=args.sigma, log=lambda s: print(s, flush=True))
    print(json.dumps(res), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
