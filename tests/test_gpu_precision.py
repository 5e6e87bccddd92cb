"""The precision-study iteration (ladder + ring, T <- R / D, E) on the GPU
against the FP64 einsum oracle, then every mixed mode against the FP64 run,
with the error stated and bounded per term.

System: o=12, v=48 (ladder GEMM 2304 x 144 x 2304, ring 576^3), occupied
eps ~ U[-2, -0.5], virtual eps ~ U[0.2, 3], integrals antisymmetrised
N(0, 0.05^2), MP2 amplitudes.  Bounds are normwise, max |dX| / max |X|, and
|dE| / |E|; each sits well above what the mode delivers at this size and far
below what a wrong digit weight, level or slice layout produces (1e-3 and
up)."""
import numpy as np
import pytest
import torch

import oracle
from dtcf_amd import mixed, precision

pytestmark = pytest.mark.gpu
FP64_TOL = 1e-12

# name, prec, nslice, level, bound on each term's R error and on |dE| / |E|
MODES = [
    ("fp32", mixed.F32, 1, 0, 5e-5),
    ("bf16x3", mixed.BF16X3, 3, 0, 5e-5),
    ("fp16x3", mixed.F16X3, 3, 0, 5e-5),
    ("int8x4/L5", mixed.I8, 4, 5, 2e-6),
    ("int8x4/L6", mixed.I8, 4, 6, 1e-7),
    ("int8x4/L6, ring L5", mixed.I8, 4, (6, 5), 2e-6),
    ("int8x5", mixed.I8, 5, 0, 5e-9),
    ("int8x6 (one GEMM per product)", mixed.I8, 6, 0, 1e-10),
]


def relerr(got, ref):
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


@pytest.fixture(scope="module")
def c4(dev):
    o, v = 12, 48
    sysm = precision.System(o, v, dev)
    host = {k: getattr(sysm, k).cpu().numpy() for k in ("Vvvvv", "Vvvoo", "Wr", "T", "eo", "ev")}
    W = np.transpose(host["Wr"].reshape(v, o, v, o), (3, 1, 2, 0))  # W[k,j,c,b] = Wr[b,j,c,k]
    ref = oracle.ccd_iteration(host["Vvvvv"].reshape(v, v, v, v), host["Vvvoo"].reshape(v, v, o, o),
                               np.ascontiguousarray(W), host["T"].reshape(v, v, o, o), host["eo"],
                               host["ev"])
    R, Tn, E, ph, parts = sysm.iterate(None, 0, 0)
    p64 = {k: x.cpu().numpy().reshape(v, v, o, o) for k, x in parts.items()}
    e64 = {k: sysm.energy_of(x) for k, x in parts.items()}
    return dict(sysm=sysm, ref=ref, Tn=Tn.cpu().numpy().reshape(v, v, o, o), E=E, parts=p64,
                eparts=e64, o=o, v=v)


def test_fp64_iteration_vs_oracle(c4):
    lad, ring, Tn, E = c4["ref"]
    assert relerr(c4["parts"]["ladder"], lad) <= FP64_TOL
    assert relerr(c4["parts"]["ring"], ring) <= FP64_TOL
    assert relerr(c4["Tn"], Tn) <= FP64_TOL
    assert abs(c4["E"] - E) <= FP64_TOL * abs(E)
    # E is linear in R: E(V_vvoo) + E(ladder) + E(ring) = E
    sysm = c4["sysm"]
    e_v = sysm.energy_of(sysm.Vvvoo)
    assert abs(e_v + c4["eparts"]["ladder"] + c4["eparts"]["ring"] - c4["E"]) <= 1e-12 * abs(c4["E"])


@pytest.mark.parametrize("mode", MODES, ids=[m[0] for m in MODES])
def test_mixed_iteration_per_term(c4, mode):
    name, prec, ns, lev, tol = mode
    sysm, o, v = c4["sysm"], c4["o"], c4["v"]
    if ns == 6:
        mixed.set_fused(False)
    try:
        R, Tn, E, ph, parts = sysm.iterate(prec, ns, lev)
    finally:
        mixed.set_fused(True)
    lines = []
    for term in ("ladder", "ring"):
        x = parts[term].cpu().numpy().reshape(v, v, o, o)
        e = relerr(x, c4["parts"][term])
        de = abs(sysm.energy_of(parts[term]) - c4["eparts"][term])
        lines.append(f"{term}: R {e:.2e}, E {de:.2e} Eh")
        assert e <= tol, (name, term, e)
        assert de <= tol * abs(c4["E"]), (name, term, de)
    dE = abs(E - c4["E"])
    print(f"{name}: " + "; ".join(lines) + f"; total E error {dE:.2e} Eh (E = {c4['E']:.6f})",
          flush=True)
    assert dE <= tol * abs(c4["E"])
    assert relerr(Tn.cpu().numpy().reshape(v, v, o, o), c4["Tn"]) <= tol
