"""The allocation-free UHF CCSD step (dtcf_uhf_ccsd_step_f64_noalloc)
against the allocating one (dtcf_uhf_ccsd_step_f64), both through the C
ABI: two calls in a row on the shared device scratch, the second with a
smaller system, each compared bit for bit (same kernels, same summation
order) on the updated amplitudes and the eight outputs; then the energy-only
form."""
import ctypes as C

import numpy as np
import pytest
import torch

from dtcf_amd import ccsd
from dtcf_amd._lib import check, lib

pytestmark = pytest.mark.gpu


def _random(Oa, Ob, Va, Vb, seed):
    rng = np.random.default_rng(seed)
    a = {"eoa": -2.0 + 1.5 * rng.random(Oa), "eob": -2.0 + 1.5 * rng.random(Ob),
         "eva": 0.2 + 2.8 * rng.random(Va), "evb": 0.2 + 2.8 * rng.random(Vb),
         "fa_ov": 0.1 * rng.standard_normal((Oa, Va)), "fb_ov": 0.1 * rng.standard_normal((Ob, Vb)),
         "vaa_oovv": 0.05 * rng.standard_normal((Oa, Oa, Va, Va)),
         "vab_oovv": 0.05 * rng.standard_normal((Oa, Ob, Va, Vb)),
         "vbb_oovv": 0.05 * rng.standard_normal((Ob, Ob, Vb, Vb))}
    shapes = {"ta_vo": (Va, Oa), "tb_vo": (Vb, Ob), "taa_vvoo": (Va, Va, Oa, Oa),
              "tab_vvoo": (Va, Vb, Oa, Ob), "tbb_vvoo": (Vb, Vb, Ob, Ob)}
    for n, shp in shapes.items():
        a[n] = 0.01 * rng.standard_normal(shp)
        a[ccsd.RESID[n]] = 0.05 * rng.standard_normal(shp)
    return a


def _run(entry, a, update, dev):
    d = {k: torch.from_numpy(np.ascontiguousarray(x)).to(dev) for k, x in a.items()}
    sizes = {"Oa": a["eoa"].size, "Ob": a["eob"].size, "Va": a["eva"].size, "Vb": a["evb"].size}
    ptr = {k: t.data_ptr() for k, t in d.items()}
    desc = ccsd.make_desc(sizes, ptr, ptr)
    out = torch.zeros(8, dtype=torch.float64, device=dev)
    f = getattr(lib(), entry)
    f.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    check(f(C.byref(desc), 1 if update else 0, C.c_void_p(out.data_ptr()),
            C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    return {k: d[k].cpu().numpy() for k in ccsd.AMPS}, out.cpu().numpy()


@pytest.mark.parametrize("update", [True, False])
def test_noalloc_matches_allocating_entry(dev, update):
    for sizes in ((12, 10, 70, 66), (3, 2, 5, 4)):  # second call smaller
        print("noalloc", sizes, "update" if update else "energy only", flush=True)
        a = _random(*sizes, seed=sum(sizes))
        t_ref, o_ref = _run("dtcf_uhf_ccsd_step_f64", a, update, dev)
        t_got, o_got = _run("dtcf_uhf_ccsd_step_f64_noalloc", a, update, dev)
        for n in ccsd.AMPS:
            assert np.array_equal(t_got[n], t_ref[n]), (sizes, n)
        assert np.array_equal(o_got, o_ref), (sizes, o_got, o_ref)
