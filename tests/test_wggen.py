"""CPU checks of the workgroup-kernel tables (pdrchem/wggen.py, pdrchem/sparse.py).

The HIP kernel (csrc/bdf_wg_impl.h) is table-driven: it executes, lane by
lane, the lists the generator emits.  These tests execute the same lists in
numpy, with the kernel's lane/round layout and level order, and compare with
independent computations:

* RHS (fluxes, chunk sums, row sums) against the oracle's pdr_rhs
  (odes_full.c:26-700 semantics);
* Jacobian entries against the oracle's bordered Jacobian, itself pinned by
  central differences of the RHS in test_oracle.py;
* the M build, level-ordered head elimination, Schur complement from the
  L_TH / U_HT position tables, tail inverse and the five-stage solve against
  numpy.linalg.solve of the dense bordered matrix;
* the bordered solve against the un-bordered Newton matrix I - gamma*(J0 + g q^T),
  which is what calculate_abundances' CVODE solve factors.
"""
import numpy as np
import pytest

import oracle
from pdrchem import sparse, workloads, wggen
from pdrchem.netgen import load_network

NETS = ["mynetwork", "full"]


@pytest.fixture(scope="module", params=NETS)
def case(request):
    name = request.param
    net = load_network(name)
    tab = wggen.WgTables(net)
    orc = oracle.Oracle(name)
    P = workloads.random_points(3, seed=7, columns=True)
    rate = orc.rates(P.T, P.Td, P.G, P.Av, P.cols)[1]
    rng = np.random.default_rng(11)
    x = 10.0 ** rng.uniform(-14, -1, net.neq)
    return name, net, tab, orc, rate, float(P.n_H[1]), x


def _kp(tab, net, rate, nH):
    kp = np.zeros(tab.nslot)
    for s, r in enumerate(tab.slot_reac):
        rr = net.reactions[r]
        kp[s] = rate[r] * (nH if rr.nh else 1.0)
    return kp


def _xvec(net, x):
    q = np.asarray(net.charges(), dtype=float)
    xe = float(np.dot(q, x))
    return np.concatenate([x, [xe, 1.0]]), xe


def _coef(v, bit):
    c = (v >> bit) & 0xF
    return c - 16 if c >= 8 else c


def test_rhs_tables(case):
    name, net, tab, orc, rate, nH, x = case
    X, xe = _xvec(net, x)
    kp = _kp(tab, net, rate, nH)
    fab = np.asarray(tab.slot_fab, dtype=np.int64)
    F = np.zeros(tab.nslot + 1)
    F[: tab.nslot] = kp * X[fab & 0xFF] * X[(fab >> 8) & 0xFF]
    P = np.zeros(tab.nv + 1)
    for p in range(tab.vl):
        for lane in range(tab.block):
            acc = 0.0
            for t in range(tab.tm):
                w = tab.rhs_term[p][t][lane]
                acc += _coef(w, 12) * F[w & 0xFFF]
            P[p * tab.block + lane] = acc
    yd = np.array([sum(P[tab.row_base[i] + u] for u in range(tab.row_nl[i])) for i in range(net.neq)])
    ref, xe_ref = orc.rhs(rate, nH, x)
    assert xe == pytest.approx(xe_ref, rel=1e-14)
    scale = np.maximum(np.abs(ref), 1e-300)
    # sums of many terms of both signs: compare against the size of the terms
    mag = np.zeros(net.neq)
    for r in net.reactions:
        if not r.stoich:
            continue
        fl = abs(rate[r.index] * (nH if r.nh else 1.0) * X[r.fa] * X[r.fb])
        for i, c in r.stoich.items():
            mag[i] += abs(c) * fl
    assert np.all(np.abs(yd - ref) <= 1e-13 * np.maximum(mag, scale)), name


def test_jacobian_tables(case):
    name, net, tab, orc, rate, nH, x = case
    X, xe = _xvec(net, x)
    kp = _kp(tab, net, rate, nH)
    fab = np.asarray(tab.slot_fab, dtype=np.int64)
    D = np.zeros(2 * tab.nslot + 1)
    D[0:2 * tab.nslot:2] = kp * X[(fab >> 8) & 0xFF]
    D[1:2 * tab.nslot:2] = kp * X[fab & 0xFF]
    Jg = np.full(tab.mtmp * tab.block, np.nan)   # the saved Jacobian, by LU position
    for lane in range(tab.block):
        acc = 0.0
        for t in range(tab.jtm):
            w = tab.jac_term[t][lane]
            acc += _coef(w, 13) * D[w & 0x1FFF]
            if (w >> 17) & 1:
                Jg[w >> 18] = acc
                acc = 0.0
    assert not np.any(np.isnan(Jg[tab.jpos])), "every entry is written"
    assert np.isnan(Jg).sum() == Jg.size - len(tab.jpos), "only the entries' positions are written"
    Jb = orc.jac_bordered(rate, nH, x)
    got = np.zeros_like(Jb)
    for e, (i, j) in enumerate(tab.jent):
        got[i, j] = Jg[tab.jpos[e]]
    mask = np.zeros_like(Jb, dtype=bool)
    for (i, j) in tab.jent:
        mask[i, j] = True
    assert np.all(Jb[~mask] == 0.0), "no oracle entry outside the table pattern"
    np.testing.assert_allclose(got, Jb, rtol=1e-12, atol=1e-300)


def _spmv(tab, sp, LU, B):
    """Chunk partials of one solve stage: P[lane] = sum over the lane's chunk."""
    P = np.zeros(tab.block)
    for rnd in sp["tab"]:
        for lane, v in enumerate(rnd):
            P[lane] += LU[v & 0x3FFF] * B[v >> 14]
    return [P[sp["base"][r]:sp["base"][r] + sp["cnt"][r]].sum() for r in range(len(sp["base"]))]


def _run_setup_stream(tab, LU):
    """The kernel's setup stream on numpy arrays: phase by phase, every lane's
    items in round order, all lanes of a round together (no item of a phase
    reads what another writes, which the generator checks).  A phase ends
    with the group whose first word carries the phase-end flag."""
    st = np.asarray(tab.stream, dtype=np.uint64).reshape(-1, tab.block)
    off = lambda v, sh: (((v >> np.uint64(sh)) & np.uint64(0xFFFC)) >> np.uint64(2)).astype(np.int64)
    bit = lambda v, sh: ((v >> np.uint64(sh)) & np.uint64(1)).astype(bool)
    r0 = 0
    for ph, ng in enumerate(tab.sp_groups):
        acc = np.zeros(tab.block)
        for g in range(ng):
            assert bool(np.all(bit(st[r0 + g * tab.SU], 0))) == (g == ng - 1), "phase-end flag"
        for rnd in st[r0:r0 + ng * tab.SU]:
            a, b, c_, t = off(rnd, 0), off(rnd, 16), off(rnd, 32), off(rnd, 48)
            last = bit(rnd, 48)
            if ph < tab.nep:
                store = bit(rnd, 49)
                m = LU[a] / LU[c_]
                acc += m * LU[b]
                new = np.where(store, m, LU[t] - acc)
                LU[t[last]] = new[last]
            else:
                acc += LU[a] * LU[b]
                LU[t[last]] = (-acc * LU[c_])[last]
            acc[last] = 0.0
        r0 += ng * tab.SU
    # the pivot check table names every head pivot's diagonal and reciprocal
    for w in tab.piv_chk:
        d, r = (w & 0xFFFF) >> 2, (w >> 16) >> 2
        assert LU[r] == 1.0 / LU[d]


def _factor_and_solve(tab, net, Jb, gamma, w, rhs):
    """The kernel's setup + solve, in the kernel's order, on numpy arrays."""
    n, NB = net.neq, tab.nb
    q = np.asarray(net.charges(), dtype=float)
    # the saved Jacobian by LU position, the bordered row's charges beside it
    Jg = np.zeros(tab.mtmp * tab.block)
    for e, (i, j) in enumerate(tab.jent):
        Jg[tab.jpos[e]] = Jb[i, j]
    LU = np.zeros(tab.nlu)
    for p in range(tab.nlu):
        d = tab.mpos[p]
        row, col = d & 0xFF, (d >> 8) & 0xFF
        if d & wggen.MF_SKIP:
            continue
        if d & wggen.MF_FIXED:
            LU[p] = 1.0 if d & wggen.MF_ONE else 0.0
            continue
        assert bool(d & wggen.MF_BORD) == (row == n) and bool(d & wggen.MF_DIAG) == (row == col)
        jv = (q[col] if row == n else Jg[p]) if d & wggen.MF_J else 0.0
        m = (-1.0 if d & wggen.MF_BORD else -gamma) * jv + (1.0 if d & wggen.MF_DIAG else 0.0)
        LU[p] = m * w[row] / w[col]
    _run_setup_stream(tab, LU)
    NT, TS, KH = tab.nt, tab.tstride, tab.kh
    T = np.array([[LU[tab.tbase + r * TS + c] for c in range(NT)] for r in range(NT)])
    if KH:
        Lth = np.array(tab.lth, dtype=np.int64).reshape(64, KH)
        Uht = np.array(tab.uht, dtype=np.int64).reshape(KH, 64)
        T = T - LU[Lth][:NT] @ LU[Uht][:, :NT]
    Tinv = np.linalg.inv(T)
    # solve: S1 forward, S2 gather, tail, S4 scatter, S5 backward
    B = np.zeros(NB + 1)
    B[:n] = rhs * w[:n]
    head, tail = tab.head, tab.tail
    s1, s2, s4, s5 = tab.spmv
    if tab.nh_piv:
        p = _spmv(tab, s1, LU, B)
        for a, v in enumerate(p):
            B[head[a]] += v
        p2 = _spmv(tab, s2, LU, B)
    else:
        p2 = [0.0] * NT
    bt = np.array([B[tail[r]] - p2[r] for r in range(NT)])
    xt = Tinv @ bt
    for r in range(NT):
        B[tail[r]] = xt[r]
    if tab.nh_piv:
        p = _spmv(tab, s4, LU, B)
        for a, v in enumerate(p):
            B[head[a]] -= v
        p = _spmv(tab, s5, LU, B)
        for a, v in enumerate(p):
            B[head[a]] = v
    return B[:n] / w[:n]


@pytest.mark.parametrize("gamma", [1e3, 1e9])
def test_factor_solve_tables(case, gamma):
    name, net, tab, orc, rate, nH, x = case
    n = net.neq
    Jb = orc.jac_bordered(rate, nH, x)
    rng = np.random.default_rng(5)
    w = np.concatenate([1.0 / (1e-8 * x + 1e-30), [1.0 / (1e-8 * abs(np.dot(net.charges(), x)) + 1e-30)]])
    rhs = rng.standard_normal(n) * x
    got = _factor_and_solve(tab, net, Jb, gamma, w, rhs)
    # dense reference: the un-bordered Newton matrix with the x_e chain term
    q = np.asarray(net.charges(), dtype=float)
    M = np.eye(n) - gamma * (Jb[:n, :n] + np.outer(Jb[:n, n], q))
    ref = np.linalg.solve(M, rhs)
    err = np.abs(got - ref) * w[:n]
    assert np.max(err) < 1e-6 * max(1.0, np.max(np.abs(ref) * w[:n])), (name, gamma, np.max(err))


def test_schur_tables(case):
    """The MFMA Schur loops' per-lane operand words (schur32, schur64), run
    the way the kernels run them, give T - L_TH U_HT of the position tables;
    in schur64 every column tile's word carries the same A operands (the
    kernel reads A from the first tile's word only)."""
    name, net, tab, *_ = case
    KH = tab.kh
    if not KH:
        assert tab.schur32 == [] and tab.schur64 == []
        return
    rng = np.random.default_rng(3)
    LU = rng.standard_normal(tab.nlu + 1)
    LU[tab.luzero] = 0.0
    Lth = np.array(tab.lth, dtype=np.int64).reshape(64, KH)
    Uht = np.array(tab.uht, dtype=np.int64).reshape(KH, 64)
    ref = -LU[Lth] @ LU[Uht]
    f = lambda w, sh: (w >> sh) & 0xFFFF
    # FP32: wave w -> tile (w >> 1, w & 1) of 32 x 32; v_mfma_f32_32x32x2f32
    ns = KH // 4
    t32 = np.array(tab.schur32, dtype=np.uint64)
    assert len(t32) == 4 * ns * 64 + wggen.SCHUR_AHEAD * wggen.SCHUR_G32 * 64
    got = np.zeros((64, 64))
    for w in range(4):
        ti, tj = w >> 1, w & 1
        words = t32[w * ns * 64:(w + 1) * ns * 64].reshape(ns, 64).astype(np.int64)
        for s in range(ns):
            for half in (0, 32):
                A = np.zeros((32, 2)); B = np.zeros((2, 32))
                for lane in range(64):
                    li, lk = lane & 31, lane >> 5
                    A[li, lk] = -LU[f(words[s, lane], half)]
                    B[lk, li] = LU[f(words[s, lane], half + 16)]
                got[32 * ti:32 * ti + 32, 32 * tj:32 * tj + 32] += A @ B
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)
    # FP64: wave w -> row band 16w, four 16 x 16 column tiles; v_mfma_f64_16x16x4f64
    ns = (KH + 7) // 8
    t64 = np.array(tab.schur64, dtype=np.uint64)
    assert len(t64) == 4 * ns * 4 * 64 + wggen.SCHUR_AHEAD * 4 * 64
    words = t64[:4 * ns * 4 * 64].reshape(4, ns, 4, 64).astype(np.int64)
    got = np.zeros((64, 64))
    for w in range(4):
        for s in range(ns):
            for jt in range(4):
                assert np.array_equal(f(words[w, s, jt], 0), f(words[w, s, 0], 0))
                assert np.array_equal(f(words[w, s, jt], 32), f(words[w, s, 0], 32))
                for half in (0, 32):
                    A = np.zeros((16, 4)); B = np.zeros((4, 16))
                    for lane in range(64):
                        li, lk = lane & 15, lane >> 4
                        A[li, lk] = -LU[f(words[w, s, jt, lane], half)]
                        B[lk, li] = LU[f(words[w, s, jt, lane], half + 16)]
                    got[16 * w:16 * w + 16, 16 * jt:16 * jt + 16] += A @ B
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_analysis_structure():
    """Structural facts the kernel relies on."""
    for name in NETS:
        net = load_network(name)
        an = sparse.analyse(net)
        assert an.nb == net.neq + 1 and an.ntail <= 64
        assert sorted(an.order) == list(range(an.nb))
        # head elimination never needs pivoting by position: every head pivot
        # is a structural nonzero
        A = sparse.bordered_pattern(net)
        assert all(A[k, k] for k in an.head)
        # the electron border stays in the tail (the head never divides by it)
        assert (an.nb - 1) in an.tail


def test_rhs_word_encoding(case):
    """the kernel's 32-bit RHS words decode to the packed terms' slot and
    coefficient (byte offset 8*slot, coefficient as an FP64 high half)"""
    import struct
    name, net, tab, orc, rate, nH, x = case
    for p in range(tab.vl):
        for t in range(tab.tm):
            for lane in range(tab.block):
                w16 = tab.rhs_term[p][t][lane]
                w = wggen.rhs_word32(w16)
                assert (w & 0xFFFF) == 8 * (w16 & 0xFFF)
                coef = struct.unpack("<d", struct.pack("<Q", (w & 0xFFFF0000) << 32))[0]
                assert coef == float(_coef(w16, 12))
