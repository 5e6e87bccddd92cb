"""DLTC tile contractions on the GPU vs the reference's checkers (checkM /
checkS restated, DLTC/checkerM.cc:39-47, checkerS.cc:35-41) and the
tile-by-tile restatement of task_mul (DLTC/task.cc:83-137)."""
import sys

import numpy as np
import pytest
import torch

import oracle
from dtcf_amd import dltc, synth

pytestmark = pytest.mark.gpu
TOL = 1e-12


def say(*a):
    print(*a, flush=True)


def relerr(got, ref):
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


def dev_t(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


# CHECK build of the reference: H1..H6 = 0..5, P1..P6 = 6..11, NOCC = 6
H = {f"H{i}": i - 1 for i in range(1, 7)}
P = {f"P{i}": 5 + i for i in range(1, 7)}


@pytest.mark.parametrize("O,V,wb", [(4, 4, 1), (4, 4, 2), (6, 8, 2)])
def test_checkM_term(dev, O, V, wb):
    # main.cc~:280-284: C[P1,P2,H1,H2] += A[P1,H4,H1,H3] * B[P2,H3,H2,H4]
    say("checkM", O, V, wb)
    rnd = oracle.Drand48()  # 2*drand48()-1 from the default seed (checkerM.cc:25-26)
    A = 2.0 * rnd.fill(V * O * O * O) - 1.0
    B = 2.0 * rnd.fill(V * O * O * O) - 1.0
    sC = oracle.checkM(A, B, O, V)
    sp = dltc.make_space(O, V, wb, wb, nocc=6)
    Cd = torch.zeros(V * V * O * O, dtype=torch.float64, device=dev)
    term = dltc.Term(sp, [P["P1"], P["P2"], H["H1"], H["H2"]], Cd,
                     [P["P1"], H["H4"], H["H1"], H["H3"]], dev_t(A, dev),
                     [P["P2"], H["H3"], H["H2"], H["H4"]], dev_t(B, dev), 1.0)
    assert term.m_ids == [H["H1"], P["P1"]] and term.n_ids == [H["H2"], P["P2"]]
    assert term.k_ids == [H["H3"], H["H4"]]
    term.run()
    got = Cd.cpu().numpy()
    assert relerr(got, sC) <= TOL
    # the reference's check_diff bound, with fabs (checker.h:15-24)
    assert np.max(np.abs(got - sC)) <= 1e-8


def test_checkS_term(dev):
    # main.cc~:286-290: X[P1,P2,H1,H2] += 1.0 * Y[P2,P1,H2,H1]
    O = V = 4
    rnd = oracle.Drand48()
    A = 2.0 * rnd.fill(V * V * O * O) - 1.0
    sC = oracle.checkS(A, O, V)
    for wb in (1, 2, 4):
        sp = dltc.make_space(O, V, wb, wb, nocc=6)
        Xd = torch.zeros(V * V * O * O, dtype=torch.float64, device=dev)
        dltc.task_sum(sp, [P["P1"], P["P2"], H["H1"], H["H2"]], Xd,
                      [P["P2"], P["P1"], H["H2"], H["H1"]], dev_t(A, dev), 1.0)
        assert np.array_equal(Xd.cpu().numpy(), sC)  # bit-exact


def _ladder_inputs(o, v, seed=11):
    # antisymmetrised N(0, 0.05^2) inputs
    V = synth.antisymmetrize(synth.normal(v ** 4, seed, 0.05).reshape(v, v, v, v))
    T = synth.antisymmetrize(synth.normal(v * v * o * o, seed + 1, 0.05).reshape(v, v, o, o))
    return V, T


@pytest.mark.parametrize("wbO,wbV", [(8, 32), (4, 8), (2, 16)])
def test_ccd_ladder_config1(dev, wbO, wbV):
    # R[a,b,i,j] += 0.5 sum_cd V[a,b,c,d] T[c,d,i,j], o=8, v=32
    # (ccsd-t2.cc:57-61: R[p3,p4,h1,h2] += 0.5 * T[p5,p6,h1,h2] * V[p3,p4,p5,p6])
    o, v = 8, 32
    say("ladder", wbO, wbV)
    V, T = _ladder_inputs(o, v)
    h = {f"h{i}": i - 1 for i in range(1, 13)}
    p = {f"p{i}": 11 + i for i in range(1, 13)}
    c_ids = [p["p3"], p["p4"], h["h1"], h["h2"]]
    a_ids = [p["p5"], p["p6"], h["h1"], h["h2"]]
    b_ids = [p["p3"], p["p4"], p["p5"], p["p6"]]
    sp = dltc.make_space(o, v, wbO, wbV)
    ref = 0.5 * np.einsum("abcd,cdij->abij", V, T)
    # expression-level task (order 6 -> SUMLEV in exec.cc, but one task covers all)
    Rd = torch.zeros(v * v * o * o, dtype=torch.float64, device=dev)
    term = dltc.Term(sp, c_ids, Rd, a_ids, dev_t(T, dev), b_ids, dev_t(V, dev), 0.5)
    term.run()
    got = Rd.cpu().numpy().reshape(v, v, o, o)
    assert relerr(got, ref) <= TOL
    # antisymmetry is preserved: R[a,b] = -R[b,a], R[i,j] = -R[j,i]
    assert np.max(np.abs(got + np.transpose(got, (1, 0, 2, 3)))) <= 1e-13 * np.max(np.abs(got))
    # sum-tile granularity: one task per (ext tile, sum tile), exec.cc:93-109
    Rd2 = torch.zeros_like(Rd)
    term2 = dltc.Term(sp, c_ids, Rd2, a_ids, dev_t(T, dev), b_ids, dev_t(V, dev), 0.5)
    spO, spV = o // wbO, v // wbV
    import itertools
    ext = [p["p3"], p["p4"], h["h1"], h["h2"]]
    sumi = [p["p5"], p["p6"]]
    ranges = [range(spV), range(spV), range(spO), range(spO)]
    ntask = 0
    for e in itertools.product(*ranges):
        for s in itertools.product(range(spV), range(spV)):
            lb = [0] * 24
            ub = [0] * 24
            for x, val in zip(ext, e):
                lb[x], ub[x] = val, val + 1
            for x, val in zip(sumi, s):
                lb[x], ub[x] = val, val + 1
            term2.task(lb, ub)
            ntask += 1
    got2 = Rd2.cpu().numpy().reshape(v, v, o, o)
    assert relerr(got2, ref) <= TOL
    say("tasks", ntask)


def test_task_graph_replays_match_eager(dev):
    """The 1024 tile tasks of the o=8, v=32 ladder (wbO=4, wbV=8) recorded as
    one HIP graph, the 64 output tiles' chains of 16 sum-tile tasks spread
    over 1 and 4 branches: every replay -- with eager kernels and
    synchronisations between replays -- gives bitwise the result of issuing
    the tasks one by one (same kernels, same order per output tile)."""
    import itertools
    o, v, wbO, wbV = 8, 32, 4, 8
    V, T = _ladder_inputs(o, v)
    h = {f"h{i}": i - 1 for i in range(1, 13)}
    p = {f"p{i}": 11 + i for i in range(1, 13)}
    c_ids = [p["p3"], p["p4"], h["h1"], h["h2"]]
    a_ids = [p["p5"], p["p6"], h["h1"], h["h2"]]
    b_ids = [p["p3"], p["p4"], p["p5"], p["p6"]]
    sp = dltc.make_space(o, v, wbO, wbV)
    Td, Vd = dev_t(T, dev), dev_t(V, dev)
    R0 = torch.from_numpy(np.random.default_rng(8).standard_normal(v * v * o * o)).to(dev)
    spO, spV = o // wbO, v // wbV
    chains = []
    for e in itertools.product(range(spV), range(spV), range(spO), range(spO)):
        boxes = []
        for sm in itertools.product(range(spV), range(spV)):
            lb, ub = [0] * 24, [0] * 24
            for x, val in zip(c_ids + [p["p5"], p["p6"]], e + sm):
                lb[x], ub[x] = val, val + 1
            boxes.append((lb, ub))
        chains.append(boxes)
    Re = R0.clone()
    eager = dltc.Term(sp, c_ids, Re, a_ids, Td, b_ids, Vd, 0.5)
    for boxes in chains:
        for lb, ub in boxes:
            eager.task(lb, ub)
    torch.cuda.synchronize()
    ref = Re.cpu()
    Rg = R0.clone()
    term = dltc.Term(sp, c_ids, Rg, a_ids, Td, b_ids, Vd, 0.5)

    def chain(boxes):
        return lambda: [term.task(lb, ub) for lb, ub in boxes]

    for branches in (1, 4):
        g = dltc.TaskGraph([chain(b) for b in chains], branches=branches)
        for rep in range(3):
            Rg.copy_(R0)
            eager.run()                      # eager work between replays
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            say("graph replay", branches, rep)
            assert torch.equal(Rg.cpu(), ref)


def test_task_graph_with_split_k_tasks(dev):
    """Tile tasks whose GEMMs split K (64 x 256 results, K = 256: four 64 x 64
    tiles on far more slots, so every tile is cut over K into a stream
    workspace) recorded as a graph on fresh streams: the workspaces are
    reserved before capture, nothing allocates under it, and every replay
    equals the eager result bitwise and the einsum to 1e-12."""
    import itertools
    o, v, wbO, wbV = 8, 32, 8, 16
    V, T = _ladder_inputs(o, v)
    h = {f"h{i}": i - 1 for i in range(1, 13)}
    p = {f"p{i}": 11 + i for i in range(1, 13)}
    c_ids = [p["p3"], p["p4"], h["h1"], h["h2"]]
    a_ids = [p["p5"], p["p6"], h["h1"], h["h2"]]
    b_ids = [p["p3"], p["p4"], p["p5"], p["p6"]]
    sp = dltc.make_space(o, v, wbO, wbV)
    Td, Vd = dev_t(T, dev), dev_t(V, dev)
    R0 = np.random.default_rng(9).standard_normal(v * v * o * o)
    spV = v // wbV
    chains = []
    for e in itertools.product(range(spV), range(spV)):
        boxes = []
        for sm in itertools.product(range(spV), range(spV)):
            lb, ub = [0] * 24, [0] * 24
            for x, val in zip([p["p3"], p["p4"], p["p5"], p["p6"]], e + sm):
                lb[x], ub[x] = val, val + 1
            lb[h["h1"]], ub[h["h1"]] = 0, 1
            lb[h["h2"]], ub[h["h2"]] = 0, 1
            boxes.append((lb, ub))
        chains.append(boxes)
    Re = dev_t(R0.copy(), dev)
    eager = dltc.Term(sp, c_ids, Re, a_ids, Td, b_ids, Vd, 0.5)
    for boxes in chains:
        for lb, ub in boxes:
            eager.task(lb, ub)
    torch.cuda.synchronize()
    ref = Re.cpu()
    full = R0.reshape(v, v, o, o) + 0.5 * np.einsum("cdij,abcd->abij", T.reshape(v, v, o, o),
                                                    V.reshape(v, v, v, v))
    assert relerr(ref.numpy().reshape(v, v, o, o), full) <= TOL
    Rg = dev_t(R0.copy(), dev)
    term = dltc.Term(sp, c_ids, Rg, a_ids, Td, b_ids, Vd, 0.5)

    def chain(boxes):
        return lambda: [term.task(lb, ub) for lb, ub in boxes]

    for lb, ub in chains[0][:1]:  # first calls set kernel attributes outside capture
        Rg.copy_(torch.from_numpy(R0))
        term.task(lb, ub)
    torch.cuda.synchronize()
    g = dltc.TaskGraph([chain(b) for b in chains], branches=4)
    for rep in range(3):
        Rg.copy_(torch.from_numpy(R0))
        eager.run()                      # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        say("split-K graph replay", rep)
        assert torch.equal(Rg.cpu(), ref)


def test_task_with_empty_index_set(dev):
    """A term whose A operand is all summed (m has no index):
    R[p2,h1] += -1 * T[p3,h4] * V[h4,p2,h1,p3] (DLTC/ccsd-t1.cc, m[0]).  Its
    tile tasks run each GEMM once: the sum over every (ext, sum) tile equals
    the whole term."""
    import itertools
    O, V, wbO, wbV = 4, 8, 2, 4
    rng = np.random.default_rng(13)
    h = {f"h{i}": i - 1 for i in range(1, 13)}
    p = {f"p{i}": 11 + i for i in range(1, 13)}
    T = rng.standard_normal((V, O))
    W = rng.standard_normal((O, V, O, V))
    R0 = rng.standard_normal((V, O))
    c_ids, a_ids, b_ids = [p["p2"], h["h1"]], [p["p3"], h["h4"]], [h["h4"], p["p2"], h["h1"], p["p3"]]
    sp = dltc.make_space(O, V, wbO, wbV)
    Rd = dev_t(R0.copy().ravel(), dev)
    term = dltc.Term(sp, c_ids, Rd, a_ids, dev_t(T, dev), b_ids, dev_t(W, dev), -1.0)
    assert term.m_ids == []
    nt = {p["p2"]: V // wbV, h["h1"]: O // wbO, p["p3"]: V // wbV, h["h4"]: O // wbO}
    for e in itertools.product(range(nt[p["p2"]]), range(nt[h["h1"]])):
        for s_ in itertools.product(range(nt[p["p3"]]), range(nt[h["h4"]])):
            lb, ub = [0] * 24, [0] * 24
            for x, val in zip([p["p2"], h["h1"], p["p3"], h["h4"]], e + s_):
                lb[x], ub[x] = val, val + 1
            term.task(lb, ub)
    got = Rd.cpu().numpy().reshape(V, O)
    ref = R0 - np.einsum("ck,kbjc->bj", T, W)
    assert relerr(got - R0, ref - R0) <= TOL


def test_task_mul_matches_tile_restatement(dev):
    # an uneven ext-tile box vs the oracle's tile-by-tile task_mul
    O, V, wbO, wbV = 6, 8, 2, 4
    rng = np.random.default_rng(4)
    h = {f"h{i}": i - 1 for i in range(1, 13)}
    p = {f"p{i}": 11 + i for i in range(1, 13)}
    T = rng.standard_normal((V, V, O, O))
    I = rng.standard_normal((O, V, O, V))
    c_ids = [p["p3"], p["p4"], h["h1"], h["h2"]]
    a_ids = [p["p3"], p["p5"], h["h1"], h["h6"]]
    b_ids = [h["h6"], p["p4"], h["h2"], p["p5"]]
    lb = [0] * 24
    ub = [O // wbO if i < 12 else V // wbV for i in range(24)]
    lb[p["p3"]], ub[p["p3"]] = 1, 2
    lb[h["h1"]], ub[h["h1"]] = 0, 2
    lb[p["p5"]], ub[p["p5"]] = 0, 1   # one sum tile of p5
    R0 = rng.standard_normal((V, V, O, O))
    ref = R0.copy()
    oracle.dltc_task_mul(c_ids, ref, a_ids, T, b_ids, I, -0.75, lb, ub, wbO, wbV)
    sp = dltc.make_space(O, V, wbO, wbV)
    Rd = dev_t(R0.copy().ravel(), dev)
    term = dltc.Term(sp, c_ids, Rd, a_ids, dev_t(T, dev), b_ids, dev_t(I, dev), -0.75)
    term.task(lb, ub)
    got = Rd.cpu().numpy().reshape(V, V, O, O)
    assert relerr(got - R0, ref - R0) <= TOL


def test_dltc_transpose_boundary(dev):
    sp = dltc.make_space(8, 12, 2, 3)
    x = ["p5", "p6", "h1", "h2"]
    y = ["h1", "h2", "p5", "p6"]
    w = [3, 3, 2, 2]
    buf = np.arange(36, dtype=np.float64)
    out = torch.zeros(36, dtype=torch.float64, device=dev)
    dltc.transpose(dev_t(buf, dev), x, out, y, sp)
    st, ref = oracle.dltc_transpose(buf, dltc.ids_of(x), dltc.ids_of(y), w)
    assert st == 0 and np.array_equal(out.cpu().numpy(), ref)
    # identity copies (the reference leaves buf_y unwritten, transp.cc:24)
    out2 = torch.zeros_like(out)
    dltc.transpose(dev_t(buf, dev), x, out2, x, sp)
    assert np.array_equal(out2.cpu().numpy(), buf)
    # rank 3 is the reference's "transpose position error"
    with pytest.raises(Exception):
        dltc.transpose(dev_t(buf[:12], dev), ["p5", "h1", "h2"], out, ["h1", "h2", "p5"], sp)
