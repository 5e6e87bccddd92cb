"""UHF CCSD iterations on the GPU: the FP64 update / energy / norm step
against the oracle (bit-exact amplitudes, energies to 1e-12), whole
iterations (T1 + T2 residual programs, then the step) against the term-list
oracle, the fixed point of the iteration against the oracle's residual
equations, and mixed-precision iterations against FP64 from the same
amplitudes."""
import numpy as np
import pytest
import torch

import oracle
from dtcf_amd import ccsd, mixed
from dtcf_amd import serve as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


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


def _device(a, dev):
    return {k: torch.from_numpy(np.ascontiguousarray(x)).to(dev) for k, x in a.items()}


def _desc(a, d):
    sizes = {"Oa": a["eoa"].size, "Ob": a["eob"].size, "Va": a["eva"].size, "Vb": a["evb"].size}
    ptr = {k: t.data_ptr() for k, t in d.items()}
    return ccsd.make_desc(sizes, ptr, ptr)


# sizes: odd occupied pairs (scalar path), even (16-byte path), several 64 x 64 tiles
SIZES = [(3, 2, 5, 4), (4, 6, 8, 10), (12, 10, 40, 36), (9, 7, 70, 66)]


@pytest.mark.parametrize("update", [True, False])
@pytest.mark.parametrize("sizes", SIZES)
def test_step_vs_oracle(dev, sizes, update):
    say("uhf step", sizes, "update" if update else "energy only")
    a = _random(*sizes, seed=sum(sizes))
    d = _device(a, dev)
    out = torch.zeros(8, dtype=torch.float64, device=dev)
    ccsd.step(_desc(a, d), update, out)
    got = out.cpu().numpy()
    t_ref, ref = oracle.uhf_ccsd_step(a, update)
    for n in ccsd.AMPS:
        assert np.array_equal(d[n].cpu().numpy(), t_ref[n]), n  # IEEE division, same order
    # energies: sums of up to 3e5 terms of either sign, summed in another order
    for k in range(5):
        assert abs(got[k] - ref[k]) <= 1e-12 * max(abs(ref[k]), 1e-300), (k, got[k], ref[k])
    for k in range(5, 8):
        assert abs(got[k] - ref[k]) <= 1e-13 * max(abs(ref[k]), 1e-300), (k, got[k], ref[k])
    # bitwise repeatable: the energy of the same amplitudes again
    out2 = torch.zeros(8, dtype=torch.float64, device=dev)
    ccsd.step(_desc(a, d), False, out2)
    out3 = torch.zeros(8, dtype=torch.float64, device=dev)
    ccsd.step(_desc(a, d), False, out3)
    assert torch.equal(out2, out3)
    if not update:
        assert torch.equal(out2[:5], out[:5])


def test_step_rejects_bad_descriptor(dev):
    a = _random(3, 2, 5, 4, seed=1)
    d = _device(a, dev)
    desc = _desc(a, d)
    desc.raa = None
    out = torch.zeros(8, dtype=torch.float64, device=dev)
    with pytest.raises(RuntimeError, match="null residuals"):
        ccsd.step(desc, True, out)
    ccsd.step(desc, False, out)  # energy only: residuals not needed
    desc.Va = 0
    with pytest.raises(RuntimeError, match="positive"):
        ccsd.step(desc, False, out)


def _oracle_inputs(system):
    arr = {n: system.host(n) for n in system.ptr}
    arr.update({k: t.cpu().numpy() for k, t in system.eps.items()})
    return arr


def _statements(tl):
    out = []
    for s in tl.statements:
        if s.kind == "mul":
            out.append((s.lhs.name, "mul", s.lhs.idx, [(t.name, t.idx) for t in s.terms]))
        else:
            out.append((s.lhs.name, "sum", s.lhs.idx, [(t.name, t.idx, t.coef) for t in s.terms]))
    return out


def _oracle_iteration(cc, arr):
    """Residuals by the term-list oracle at arr's amplitudes, then the
    oracle's update; returns (new arr, outputs)."""
    res = {}
    for tl in (cc.t1, cc.t2):
        r = oracle.termlist_run(_statements(tl), {n: arr[n] for n in tl.inputs()})
        res.update({n: r[n] for n in tl.outputs()})
    t, out = oracle.uhf_ccsd_step({**arr, **res}, True)
    new = dict(arr)
    new.update(t)
    return new, out, res


def test_iterations_vs_oracle(dev):
    sizes = {"Oa": 3, "Ob": 2, "Va": 5, "Vb": 4}
    with S.Server(1 << 30, nslots=4) as srv:
        system = ccsd.CCSystem(srv, 0, 0, sizes=sizes, seed=3)
        cc = ccsd.UHFCCSD(srv, system)
        arr = _oracle_inputs(system)
        # Fock diagonals are the orbital energies; v_vvoo = v_oovv^T
        assert np.array_equal(np.diag(arr["fa_oo"]), arr["eoa"])
        assert np.array_equal(arr["vab_vvoo"], arr["vab_oovv"].transpose(2, 3, 0, 1))
        assert np.array_equal(arr["fb_vo"], arr["fb_ov"].T)
        for it in range(3):
            rec = cc.iterate()
            arr, ref, res = _oracle_iteration(cc, arr)
            got = cc.amplitudes()
            say(f"iteration {it}: E {rec['E']:.12f} (oracle {ref[0]:.12f}), |R2| {rec['r2']:.3e}")
            if it == 0:
                # from zero amplitudes: first-order amplitudes, bit for bit
                D1 = arr["eoa"][None, :] - arr["eva"][:, None]
                assert np.array_equal(got["ta_vo"], arr["fa_vo"] / D1)
                assert np.array_equal(got["ta_vo"], arr["ta_vo"])
                assert np.array_equal(got["tab_vvoo"], arr["tab_vvoo"])
            for n in ccsd.AMPS:
                e = np.max(np.abs(got[n] - arr[n])) / max(np.max(np.abs(arr[n])), 1e-300)
                assert e <= 1e-12, (it, n, e)
            for k, name in enumerate(ccsd.OUT_NAMES):
                assert abs(rec[name] - ref[k]) <= 1e-11 * max(abs(ref[k]), 1e-12), (it, name)
            # keep the oracle on the GPU's amplitudes (no drift between the two)
            arr.update(got)
        cc.destroy()


def test_iteration_reaches_the_residual_equations_root(dev):
    """Small integrals: the Jacobi iteration converges, and at its fixed
    point the oracle's residuals (the reference term lists evaluated by
    einsum) vanish -- the update has the sign and denominators that make
    R = 0 the fixed point."""
    sizes = {"Oa": 3, "Ob": 3, "Va": 6, "Vb": 6}
    with S.Server(1 << 30, nslots=4) as srv:
        system = ccsd.CCSystem(srv, 0, 0, sizes=sizes, seed=4, sigma=0.02, fock_sigma=0.005)
        cc = ccsd.UHFCCSD(srv, system)
        hist = []
        for it in range(60):
            rec = cc.iterate()
            hist.append(rec)
            if rec["r1"] + rec["r2"] < 1e-12:
                break
        say(f"converged after {len(hist)} iterations: E {hist[-1]['E']:.14f}, "
            f"|R1| {hist[-1]['r1']:.2e} |R2| {hist[-1]['r2']:.2e}")
        assert hist[-1]["r1"] + hist[-1]["r2"] < 1e-12
        assert abs(hist[-1]["E"] - hist[-2]["E"]) < 1e-13
        arr = _oracle_inputs(system)
        _, _, res = _oracle_iteration(cc, arr)
        worst = max(np.max(np.abs(r)) for r in res.values())
        assert worst < 1e-11, worst
        cc.destroy()


@pytest.mark.parametrize("prec,ns,lev,tol", [(mixed.I8, 5, 0, 5e-8), (mixed.I8, 4, 6, 5e-7)])
def test_mixed_iteration_against_fp64_same_state(dev, prec, ns, lev, tol):
    """From the same amplitudes, a mixed iteration (every contraction on
    INT8 digit slices, integrals split once and kept) against the FP64 one:
    each residual block and the energy, per component."""
    sizes = {"Oa": 4, "Ob": 4, "Va": 12, "Vb": 10}
    with S.Server(2 << 30, nslots=4) as srv:
        system = ccsd.CCSystem(srv, 0, 0, sizes=sizes, seed=5)
        c64 = ccsd.UHFCCSD(srv, system)
        cmx = ccsd.UHFCCSD(srv, system, prec=prec, nslice=ns, max_level=lev, mixed_min=0)
        c64.iterate()  # first-order amplitudes: every term contributes afterwards
        for it in range(2):
            start = c64.amplitudes()
            c64.residuals()
            cmx.residuals()
            r64, rmx = c64.residual_arrays(), cmx.residual_arrays()
            for n in r64:
                e = np.max(np.abs(rmx[n] - r64[n])) / np.max(np.abs(r64[n]))
                assert e <= tol, (it, n, e)
            emx = cmx.step(True)
            for n, x in start.items():
                S.from_host(system.ptr[n], x)
            e64 = c64.step(True)
            errs = {k: abs(emx[k] - e64[k]) for k in ("E", "E1", "Eaa", "Eab", "Ebb")}
            say(f"{mixed.NAMES[prec]}x{ns} L{lev} iteration {it}: E {e64['E']:.10f}, |dE| " +
                ", ".join(f"{k} {v:.1e}" for k, v in errs.items()))
            for k, v in errs.items():
                assert v <= tol * abs(e64["E"]), (it, k, v)
        c64.destroy()
        cmx.destroy()
