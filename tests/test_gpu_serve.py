"""Serving runtime on the GPU: every statement of the UHF T1/T2 lists against
the einsum oracle, mixed precision against FP64, bitwise identity of a
request's results alone vs in a mix, and budget admission."""
import numpy as np
import pytest
import torch

import oracle
from dtcf_amd import mixed, ops, termlist
from dtcf_amd import serve as S
from dtcf_amd._lib import lib

pytestmark = pytest.mark.gpu

ASYM = {"Oa": 3, "Ob": 2, "Va": 5, "Vb": 4}


@pytest.fixture(scope="module")
def lists():
    return {"t1": termlist.load("t1"), "t2": termlist.load("t2")}


def oracle_statements(tl):
    out = []
    for s in tl.statements:
        if s.kind == "mul":
            out.append((s.lhs.name, "mul", s.lhs.idx, [(t.name, t.idx) for t in s.terms]))
        else:
            out.append((s.lhs.name, "sum", s.lhs.idx, [(t.name, t.idx, t.coef) for t in s.terms]))
    return out


def oracle_run(tl, system):
    arrays = {n: system.host(n) for n in tl.inputs()}
    return oracle.termlist_run(oracle_statements(tl), arrays)


@pytest.mark.parametrize("which", ["t1", "t2"])
def test_every_statement_matches_oracle(dev, lists, which):
    tl = lists[which]
    print(f"serve parity {which} sizes {ASYM}", flush=True)
    with S.Server(1 << 30, nslots=4) as srv:
        sysm = S.System(srv, 0, 0, [tl], sizes=ASYM)
        prog = S.Program(srv, tl, sysm, expose=True)
        assert len(prog.outputs) == len(tl.statements)
        got = prog.run()
        ref = oracle_run(tl, sysm)
        worst = 0.0
        for name in prog.outputs:
            r = ref[name]
            scale = max(np.abs(r).max(), 1e-300)
            err = np.abs(got[name] - r).max() / scale
            worst = max(worst, err)
            assert err <= 1e-12, (name, err)
        print(f"  {len(prog.outputs)} arrays, worst relative error {worst:.2e}")
        prog.destroy()


@pytest.mark.parametrize("which", ["t1", "t2"])
def test_every_statement_matches_oracle_thin_kernel(dev, lists, which):
    """Every statement with the thin streaming kernel forced for the FP64
    contractions it takes (one side at most 48, K a multiple of 8: with
    these extents most contractions over an o-v or v-v pair), C written
    fresh (overwrite) through the statement's views, against the oracle."""
    tl = lists[which]
    sizes = {"Oa": 4, "Ob": 2, "Va": 8, "Vb": 6}
    print(f"serve parity thin {which} sizes {sizes}", flush=True)
    ops.set_gemm_config(6)
    try:
        with S.Server(1 << 30, nslots=4) as srv:
            sysm = S.System(srv, 0, 0, [tl], sizes=sizes)
            prog = S.Program(srv, tl, sysm, expose=True)
            got = prog.run()
            ref = oracle_run(tl, sysm)
            for name in prog.outputs:
                r = ref[name]
                err = np.abs(got[name] - r).max() / max(np.abs(r).max(), 1e-300)
                assert err <= 1e-12, (name, err)
            prog.destroy()
    finally:
        ops.set_gemm_config(-1)


@pytest.mark.parametrize("layout", [S.LAYOUT_REUSE, S.LAYOUT_SEPARATE])
@pytest.mark.parametrize("which", ["t1", "t2"])
def test_concurrent_program_matches_oracle(dev, lists, which, layout):
    """With 8 slots statements finish out of program order.  Reuse layout:
    temporaries share bytes of the request block, so a reused region must
    wait for every reader of its previous occupant.  Separate layout: only
    the data dependencies order the statements."""
    tl = lists[which]
    sizes = {"Oa": 4, "Ob": 3, "Va": 12, "Vb": 10}
    print(f"serve concurrent {which} sizes {sizes} layout {layout}", flush=True)
    with S.Server(1 << 30, nslots=8) as srv:
        sysm = S.System(srv, 0, 0, [tl], sizes=sizes)
        prog = S.Program(srv, tl, sysm, layout=layout)
        assert prog.layout == layout
        ref = oracle_run(tl, sysm)
        for rep in range(3):
            got = prog.run()
            for name in prog.outputs:
                err = np.abs(got[name] - ref[name]).max() / np.abs(ref[name]).max()
                assert err <= 1e-12, (rep, name, err)
        prog.destroy()


@pytest.mark.parametrize("prec,ns,lev,tol", [(mixed.I8, 5, 0, 5e-8), (mixed.I8, 4, 6, 5e-7),
                                             (mixed.F32, 1, 0, 5e-5), (mixed.BF16X3, 3, 0, 5e-5)])
def test_mixed_program_against_fp64(dev, lists, prec, ns, lev, tol):
    tl = lists["t2"]
    sizes = {"Oa": 4, "Ob": 4, "Va": 12, "Vb": 10}
    print(f"serve mixed prec={prec} ns={ns}", flush=True)
    with S.Server(1 << 30, nslots=4) as srv:
        sysm = S.System(srv, 0, 0, [tl], sizes=sizes)
        p64 = S.Program(srv, tl, sysm)
        pmx = S.Program(srv, tl, sysm, prec=prec, nslice=ns, max_level=lev,
                        mixed_min=0)  # every contraction
        r64, rmx = p64.run(), pmx.run()
        for name in p64.outputs:
            err = np.abs(rmx[name] - r64[name]).max() / np.abs(r64[name]).max()
            print(f"  {name}: max |dR| / max |R| = {err:.2e}")
            assert err <= tol, (name, err)
        p64.destroy()
        pmx.destroy()


def test_mixed_min_policy(dev, lists):
    """A mixed program takes FP64 for contractions with m n / (m + n) below
    mixed_min: at the default (256) every contraction of a small system is
    one, and its results equal the FP64 program's bitwise; with mixed_min 0
    they take INT8 digit slices (and differ, within the INT8x5 bound)."""
    tl = lists["t2"]
    sizes = {"Oa": 4, "Ob": 4, "Va": 12, "Vb": 10}
    with S.Server(1 << 30, nslots=2) as srv:
        sysm = S.System(srv, 0, 0, [tl], sizes=sizes)
        p64 = S.Program(srv, tl, sysm)
        pdef = S.Program(srv, tl, sysm, prec=mixed.I8, nslice=5)
        pall = S.Program(srv, tl, sysm, prec=mixed.I8, nslice=5, mixed_min=0)
        r64, rdef, rall = p64.run(), pdef.run(), pall.run()
        for name in p64.outputs:
            assert np.array_equal(rdef[name], r64[name]), name
        assert any(not np.array_equal(rall[n], r64[n]) for n in p64.outputs)
        for p in (p64, pdef, pall):
            p.destroy()


def test_resident_slices_bitwise(dev, lists):
    """Integral inputs split once at finalize give bitwise the results of
    splitting them in every request (same kernels, same digits), twice in a
    row; the slices count as resident and are released with the program.
    The amplitudes are never resident: they change every iteration."""
    tl = lists["t2"]
    sizes = {"Oa": 4, "Ob": 4, "Va": 12, "Vb": 10}
    with S.Server(1 << 30, nslots=2) as srv:
        sysm = S.System(srv, 0, 0, [tl], sizes=sizes)
        base = srv.usage()["resident"]
        res = S.integral_inputs(tl)
        assert res and not any(n.startswith("t") for n in res)
        assert len(res) < len(tl.inputs())
        p0 = S.Program(srv, tl, sysm, prec=mixed.I8, nslice=5, mixed_min=0)
        p1 = S.Program(srv, tl, sysm, prec=mixed.I8, nslice=5, mixed_min=0, resident=res)
        assert srv.usage()["resident"] > base
        r0 = p0.run()
        for _ in range(2):
            r1 = p1.run()
            for name in p0.outputs:
                assert np.array_equal(r0[name], r1[name]), name
        p0.destroy()
        p1.destroy()
        assert srv.usage()["resident"] == base


def test_rebind_after_finalize(dev, lists):
    """An input with resident slices cannot be rebound after finalize (its
    slices were cut from the old contents); another input can, between
    requests, and the program then reads the new array: its results equal a
    program finalized on the new array."""
    import ctypes as C

    tl = lists["t2"]
    sizes = {"Oa": 4, "Ob": 4, "Va": 12, "Vb": 10}
    with S.Server(1 << 30, nslots=2) as srv:
        sysm = S.System(srv, 0, 0, [tl], sizes=sizes, seed=5)
        other = S.System(srv, 0, 0, [tl], sizes=sizes, seed=6)
        res = S.integral_inputs(tl)
        p = S.Program(srv, tl, sysm, prec=mixed.I8, nslice=5, mixed_min=0, resident=res)
        p.run()
        L = lib()
        with pytest.raises(RuntimeError, match="resident slices"):
            from dtcf_amd._lib import check
            check(L.dtcf_program_bind(p.h, res[0].encode(), C.c_void_p(other.ptr[res[0]])))
        amp = "taa_vvoo"
        assert amp in tl.inputs() and amp not in res
        check(L.dtcf_program_bind(p.h, amp.encode(), C.c_void_p(other.ptr[amp])))
        got = p.run()
        # reference: the same program over sysm's inputs with other's taa_vvoo
        sysm2 = S.System(srv, 0, 0, [tl], sizes=sizes, seed=5)
        S.from_host(sysm2.ptr[amp], other.host(amp))
        pr = S.Program(srv, tl, sysm2, prec=mixed.I8, nslice=5, mixed_min=0, resident=res)
        ref = pr.run()
        for name in got:
            assert np.array_equal(got[name], ref[name]), name
        p.destroy()
        pr.destroy()


def _outputs_host(outs):
    return [o.cpu().numpy().copy() for o in outs]


def test_alone_vs_mix_bitwise(dev, lists):
    t1, t2 = lists["t1"], lists["t2"]
    sa = {"Oa": 6, "Ob": 5, "Va": 20, "Vb": 18}
    sb = {"Oa": 4, "Ob": 4, "Va": 14, "Vb": 14}
    print("serve alone-vs-mix", flush=True)
    with S.Server(4 << 30, nslots=6) as srv:
        A = S.System(srv, 0, 0, [t1, t2], sizes=sa, seed=11)
        B = S.System(srv, 0, 0, [t1, t2], sizes=sb, seed=12)
        progs = {
            "A2": S.Program(srv, t2, A), "A1": S.Program(srv, t1, A),
            "B2": S.Program(srv, t2, B), "B1": S.Program(srv, t1, B),
            "A2i8": S.Program(srv, t2, A, prec=mixed.I8, nslice=5, mixed_min=0),
        }
        torch.cuda.synchronize()
        alone = {}
        for k in ("A2", "A2i8", "B1"):
            rid, outs = progs[k].submit()
            assert srv.wait(rid, 120000) == 0
            srv.release(rid)
            alone[k] = _outputs_host(outs)
        # the same requests again, interleaved with others
        order = ["B2", "A2", "A1", "B1", "A2i8", "B2", "A1", "A2", "B1", "A2i8"]
        subs = [(k,) + progs[k].submit() for k in order]
        for k, rid, outs in subs:
            assert srv.wait(rid, 120000) == 0
        for k, rid, outs in subs:
            if k in alone:
                for x, y in zip(_outputs_host(outs), alone[k]):
                    assert np.array_equal(x, y), k
            srv.release(rid)
        for p in progs.values():
            p.destroy()


def test_results_independent_of_slot_count(dev, lists):
    tl = lists["t2"]
    sizes = {"Oa": 5, "Ob": 4, "Va": 16, "Vb": 12}
    res = []
    for nslots in (1, 8):
        with S.Server(1 << 30, nslots=nslots) as srv:
            sysm = S.System(srv, 0, 0, [tl], sizes=sizes, seed=3)
            p = S.Program(srv, tl, sysm)
            res.append(p.run())
            p.destroy()
    for n in res[0]:
        assert np.array_equal(res[0][n], res[1][n]), n


def test_budget_admission_and_rejection(dev, lists):
    tl = lists["t2"]
    sizes = {"Oa": 4, "Ob": 4, "Va": 12, "Vb": 12}
    with S.Server(1 << 30, nslots=4) as probe:
        ps = S.System(probe, 0, 0, [tl], sizes=sizes)
        pp = S.Program(probe, tl, ps)
        block, resident = pp.block_bytes, probe.usage()["resident"]
        pp.destroy()
    # room for the inputs and one and a half request blocks
    budget = resident + 4096 + block * 3 // 2
    with S.Server(budget, nslots=4) as srv:
        sysm = S.System(srv, 0, 0, [tl], sizes=sizes)
        p = S.Program(srv, tl, sysm)
        torch.cuda.synchronize()
        subs = [p.submit() for _ in range(4)]
        for rid, _ in subs:
            assert srv.wait(rid, 120000) == 0
        u = srv.usage()
        assert u["peak_reserved"] <= ((block + 255) // 256) * 256
        infos = [srv.info(rid) for rid, _ in subs]
        # one at a time: each admitted after the previous finished
        for a, b in zip(infos, infos[1:]):
            assert b.t_admit >= a.t_done
        first = [x.cpu().numpy() for x in subs[0][1]]
        for _, outs in subs[1:]:
            for x, y in zip(outs, first):
                assert np.array_equal(x.cpu().numpy(), y)
        for rid, _ in subs:
            srv.release(rid)
        p.destroy()
    # a request block that cannot fit beside the resident inputs is refused
    with S.Server(resident + 4096 + block // 2, nslots=2) as srv:
        sysm = S.System(srv, 0, 0, [tl], sizes=sizes)
        p = S.Program(srv, tl, sysm)
        with pytest.raises(RuntimeError, match="budget"):
            p.submit()
        p.destroy()


def test_task_timestamps_ordered(dev, lists):
    tl = lists["t1"]
    with S.Server(1 << 28, nslots=3) as srv:
        sysm = S.System(srv, 0, 0, [tl], sizes=ASYM)
        p = S.Program(srv, tl, sysm)
        torch.cuda.synchronize()
        rid, _ = p.submit()
        assert srv.wait(rid, 120000) == 0
        inf = srv.info(rid)
        assert inf.state == 2 and inf.ntasks == len(tl.statements)
        assert inf.t_submit <= inf.t_admit <= inf.t_done
        r, i, d = srv.task_times(rid, inf.ntasks)
        assert np.all(r >= inf.t_admit) and np.all(r <= i) and np.all(i <= d)
        assert np.all(d <= inf.t_done)
        srv.release(rid)
        p.destroy()


def test_auto_layout_choice(dev, lists):
    """AUTO keeps temporaries apart when that costs little block memory and
    reuses bytes otherwise; the separate block is never smaller."""
    tl = lists["t2"]
    with S.Server(8 << 30, nslots=4) as srv:
        sysm = S.System(srv, 0, 0, [tl], sizes={"Oa": 4, "Ob": 3, "Va": 12, "Vb": 10})
        auto = S.Program(srv, tl, sysm)
        reuse = S.Program(srv, tl, sysm, layout=S.LAYOUT_REUSE)
        sep = S.Program(srv, tl, sysm, layout=S.LAYOUT_SEPARATE)
        assert sep.block_bytes >= reuse.block_bytes
        # a tiny system: the separate block is far below the 256 MiB slack
        assert auto.layout == S.LAYOUT_SEPARATE and auto.block_bytes == sep.block_bytes
        for p in (auto, reuse, sep):
            p.destroy()


@pytest.mark.parametrize("nslots", [1, 8])
def test_failed_statement_fails_its_request_only(dev, lists, nslots):
    """A statement that fails at issue (forced on the host: nothing faults)
    fails its request -- dtcf_request_wait reports DTCF_ERR_HIP with the
    message, the request is marked failed and its block is released --
    while a request submitted beside it completes with the result it has
    alone, and the server serves the program again afterwards."""
    tl = lists["t2"]
    sizes = {"Oa": 4, "Ob": 3, "Va": 10, "Vb": 8}
    with S.Server(2 << 30, nslots=nslots) as srv:
        sysm = S.System(srv, 0, 0, [tl], sizes=sizes, seed=21)
        p = S.Program(srv, tl, sysm)
        good = p.run()
        torch.cuda.synchronize()
        srv.inject_failure(40)  # the 41st statement issued from now on
        rid_bad, _ = p.submit()
        st = srv.wait(rid_bad, 120000)
        assert st == -3, st  # DTCF_ERR_HIP
        assert b"injected" in lib().dtcf_last_error()
        inf = srv.info(rid_bad)
        assert inf.state == 3
        srv.release(rid_bad)
        assert srv.usage()["reserved"] == 0
        # a failure in a mix: the other request is unaffected
        srv.inject_failure(25)
        rid_a, outs_a = p.submit()
        rid_b, outs_b = p.submit()
        sts = [srv.wait(r, 120000) for r in (rid_a, rid_b)]
        assert sorted(sts) == [-3, 0], sts
        ok_outs = outs_a if sts[0] == 0 else outs_b
        for name, o, shp in zip(p.outputs, ok_outs, p.out_shapes):
            assert np.array_equal(o.cpu().numpy().reshape(shp), good[name]), name
        for r in (rid_a, rid_b):
            srv.release(r)
        assert srv.usage()["reserved"] == 0
        again = p.run()
        for name in good:
            assert np.array_equal(again[name], good[name]), name
        p.destroy()


def test_statement_plans(dev, lists):
    """dtcf_program_statement_plan reports, for every statement of a
    finalized T2 program, its kind, a GEMM shape whose m n k equals the
    statement's external-by-external-by-contracted extents (either operand
    order), a known tile configuration (FP64), and the mixed flag exactly on
    the contractions the mixed-precision rule selects."""
    tl = lists["t2"]
    with S.Server(1 << 30, nslots=2) as srv:
        sysm = S.System(srv, 0, 0, [tl], sizes=ASYM)
        # mixed_min 2: the rule m n >= 2 (m + n) splits most contractions at these extents
        progs = {0: S.Program(srv, tl, sysm),
                 1: S.Program(srv, tl, sysm, prec=mixed.I8, nslice=4, mixed_min=2.0)}
        for prec, prog in progs.items():
            for i, s in enumerate(tl.statements):
                pl = prog.plan(i)
                assert pl["kind"] == s.kind
                if s.kind == "sum":
                    assert pl["nterms"] == len(s.terms)
                    continue
                a, b = s.terms
                ext_a = [x for x in a.idx if x in s.lhs.idx]
                ext_b = [x for x in b.idx if x in s.lhs.idx]
                con = [x for x in a.idx if x in b.idx and x not in s.lhs.idx]
                ea, eb, ek = (int(np.prod([s.extent(x, sysm.sizes) for x in g])) for g in (ext_a, ext_b, con))
                assert pl["k"] == ek and sorted((pl["m"], pl["n"])) == sorted((ea, eb)), (s.lhs.name, pl)
                mixed_rule = prec != 0 and pl["m"] * pl["n"] >= 2 * (pl["m"] + pl["n"])
                assert ("mixed" in pl["flags"]) == mixed_rule, (s.lhs.name, pl)
                assert pl["tile"] in (("mixed",) if mixed_rule else ("128x128", "64x64", "192x64", "64x192", "thin")), pl
            if prec:
                assert any("mixed" in prog.plan(i)["flags"] for i in range(len(tl.statements)))
            prog.destroy()
        with pytest.raises(Exception):
            progs[0].plan(0)  # destroyed
