"""CPU checks of the wavefront-kernel tables (pdrchem/gpugen.py) and of the
kernel's Gauss-Jordan scheme (csrc/bdf_wave_impl.h).

The reduced-network kernel runs one system per wave64: lane L holds row
i = L & 31, columns 16*(L >> 5) .. +15, and executes the generated lists
lane by lane.  These tests run the same lists in numpy with the kernel's
lane layout and compare with independent computations:

* RHS (flux slots, per-lane term chunks, row partial sums) against the
  oracle's pdr_rhs (odes_reduced.c:26-148 semantics);
* Jacobian (flux partials, per-column term lists, x_e chain column) against
  the oracle's exact Jacobian, itself pinned by central differences of the
  RHS in test_oracle.py;
* the scaled Gauss-Jordan inverse with the kernel's packed-key pivot search
  (FP32 value bits with the row in the low five bits), kept in pivot order,
  and the mat-vec that applies it, against numpy.linalg.solve.
"""
import numpy as np
import pytest

import oracle
from pdrchem import gpugen, workloads
from pdrchem.netgen import load_network


@pytest.fixture(scope="module")
def case():
    net = load_network("reduced")
    tab = gpugen.WaveTables(net)
    orc = oracle.Oracle("reduced")
    P = workloads.random_points(3, seed=5)
    rate = orc.rates(P.T, P.Td, P.G, P.Av, P.cols)[2]
    rng = np.random.default_rng(3)
    x = 10.0 ** rng.uniform(-14, -1, net.neq)
    return net, tab, orc, rate, float(P.n_H[2]), x


def _factor_vector(net, x):
    X = np.zeros(64)
    X[: net.neq] = x
    X[net.neq] = float(np.dot(np.asarray(net.charges(), dtype=float), x))
    X[net.neq + 1] = 1.0
    return X


def _kp(tab, net, rate, nH):
    kp = np.zeros(tab.nslot)
    for s in range(tab.nslot):
        if s < len(tab.slot_reac):
            k = rate[tab.slot_reac[s]]
            kp[s] = k * nH if (tab.slot_fab[s] >> 16) & 1 else k
    return kp


def test_slot_layout(case):
    net, tab, _, _, _, _ = case
    assert tab.nslot % 64 == 0 and tab.zero_slot == tab.nslot and tab.jzero == 2 * tab.nslot
    # every reaction with a net effect owns exactly one slot
    assert sorted(tab.slot_reac) == [r.index for r in net.reactions if r.stoich]
    # padded slots multiply the constant factor by itself with rate 0
    for s in range(len(tab.slot_reac), tab.nslot):
        assert tab.slot_fab[s] == (net.neq + 1) | ((net.neq + 1) << 8)
    # RHS terms index fluxes or the zero slot, Jacobian terms partials or their zero
    assert max(s for row in tab.rhs_term for (s, _) in row) <= tab.nslot
    assert max(d for row in tab.jac_term for (d, _) in row) <= 2 * tab.nslot


def test_rhs_tables(case):
    net, tab, orc, rate, nH, x = case
    X = _factor_vector(net, x)
    kp = _kp(tab, net, rate, nH)
    fab = np.asarray(tab.slot_fab, dtype=np.int64)
    F = np.zeros(tab.nslot + 1)                     # F[nslot] = 0
    F[: tab.nslot] = kp * X[fab & 0xFF] * X[(fab >> 8) & 0xFF]
    P = np.zeros(65)                                # P[64] = 0
    for lane in range(64):
        P[lane] = sum(c * F[s] for (s, c) in (tab.rhs_term[t][lane] for t in range(tab.tm)))
    yd = np.zeros(net.neq)
    for i in range(net.neq):
        base, nl = tab.row_base[i], tab.row_nlanes[i]
        yd[i] = sum(P[min(base + u, 64)] for u in range(nl))
    ref, xe = orc.rhs(rate, nH, x)
    assert X[net.neq] == pytest.approx(xe, rel=1e-14)
    mag = np.zeros(net.neq)
    for r in net.reactions:
        fl = abs(rate[r.index] * (nH if r.nh else 1.0) * X[r.fa] * X[r.fb])
        for i, c in r.stoich.items():
            mag[i] += abs(c) * fl
    assert np.all(np.abs(yd - ref) <= 1e-13 * np.maximum(mag, np.abs(ref)))


def test_jacobian_tables(case):
    net, tab, orc, rate, nH, x = case
    X = _factor_vector(net, x)
    kp = _kp(tab, net, rate, nH)
    fab = np.asarray(tab.slot_fab, dtype=np.int64)
    D = np.zeros(2 * tab.nslot + 1)                 # D[2*nslot] = 0
    D[0:2 * tab.nslot:2] = kp * X[(fab >> 8) & 0xFF]
    D[1:2 * tab.nslot:2] = kp * X[fab & 0xFF]
    J = np.zeros((32, 32))
    off = 0
    acc = np.zeros((64, 17))
    for j in range(17):
        for t in range(tab.jt[j]):
            for lane in range(64):
                d, c = tab.jac_term[off + t][lane]
                acc[lane, j] += c * D[d]
        off += tab.jt[j]
    q = tab.charges
    for lane in range(64):
        i, hh = lane & 31, lane >> 5
        for j in range(16):
            J[i, 16 * hh + j] = acc[lane, j] + acc[lane, 16] * q[16 * hh + j]
    ref = orc.jac(rate, nH, x)
    n = net.neq
    np.testing.assert_allclose(J[:n, :n], ref, rtol=1e-12, atol=1e-300)
    assert np.all(J[n:, :] == 0.0) and np.all(J[:, n:] == 0.0)


def _pivot_key(a, i):
    """the kernel's pivot_key<5>: the leading word of |a| (FP32: its bits,
    FP64: the high word) with 31 - row in the low five bits"""
    if isinstance(a, np.float32):
        word = int(np.float32(abs(a)).view(np.uint32))
    else:
        word = int(np.float64(abs(a)).view(np.uint64)) >> 32
    return (word & ~31) | (31 - i)


def _gauss_jordan(M, dtype):
    """Gauss-Jordan on the 32x32 matrix as the kernel runs it, in the matrix
    type: rows are never moved; returns the in-place inverse in pivot order
    and invp (the step at which each row was the pivot)."""
    A = M.astype(dtype).copy()
    used = np.zeros(32, dtype=bool)
    invp = np.arange(32)
    lim = 0x7F800000 if dtype == np.float32 else 0x7FF00000
    for k in range(32):
        keys = [0 if used[i] else _pivot_key(A[i, k], i) for i in range(32)]
        kmax = max(keys)
        if kmax <= 31 or kmax >= lim:
            raise np.linalg.LinAlgError("singular")
        p = 31 - (kmax & 31)
        d = dtype(1.0) / A[p, k]
        pr = A[p].copy() * d
        fi = A[:, k].copy()
        for i in range(32):
            if i == p:
                A[i] = pr
                A[i, k] = d
            else:
                A[i] = A[i] - fi[i] * pr
                A[i, k] = -fi[i] * d
        used[p] = True
        invp[p] = k
    return A, invp


def _inv_apply(A, invp, b):
    """the kernel's inv_apply: V[invp[i]] = b_i, then row i's product with V
    is x[invp[i]]"""
    V = np.zeros(32)
    V[invp] = b
    x = np.zeros(32)
    x[invp] = A.astype(np.float64) @ V
    return x


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-14), (np.float32, 1e-6)])
def test_gauss_jordan_scheme(case, dtype, tol):
    net, tab, orc, rate, nH, x = case
    n = net.neq
    J = np.zeros((32, 32))
    J[:n, :n] = orc.jac(rate, nH, x)
    rng = np.random.default_rng(9)
    w = 1.0 / (1e-8 * x + 1e-30)                    # error weights of the state
    for gamma in (1e-6, 1e2, 1e6):
        M = np.eye(32) - gamma * J
        W = np.ones(32)
        W[:n] = w
        Ms = M * W[:, None] / W[None, :]            # W M W^-1, as the kernel scales it
        A, invp = _gauss_jordan(Ms, dtype)
        assert sorted(invp) == list(range(32))
        b = np.zeros(32)
        b[:n] = rng.standard_normal(n)
        xs = _inv_apply(A, invp, b)
        # normwise backward error of the scaled solve (the scaled matrix is
        # ill-conditioned for large gamma; the Newton iteration absorbs that)
        be = np.max(np.abs(Ms @ xs - b)) / (np.max(np.abs(Ms).sum(axis=1)) * np.max(np.abs(xs)) + np.max(np.abs(b)))
        assert be < tol, (gamma, be)
        if gamma < 1.0:                             # near the identity: forward error too
            ref = np.linalg.solve(Ms, b)
            assert np.max(np.abs(xs - ref)) < 100 * np.finfo(dtype).eps * np.max(np.abs(ref))


def test_gauss_jordan_random_matrices():
    """the pivot-order inverse on general matrices (pivoting exercised)"""
    rng = np.random.default_rng(4)
    for _ in range(5):
        R = rng.standard_normal((32, 32))
        A, invp = _gauss_jordan(R, np.float64)
        b = rng.standard_normal(32)
        np.testing.assert_allclose(_inv_apply(A, invp, b), np.linalg.solve(R, b), rtol=1e-9, atol=1e-9)
    with pytest.raises(np.linalg.LinAlgError):
        S = rng.standard_normal((32, 32))
        S[:, 5] = 0.0
        _gauss_jordan(S, np.float64)


def test_pivot_key_order():
    vals = np.float32([0.0, 1e-30, 0.5, 0.5, 3.0, 2.9999998])
    keys = [_pivot_key(v, i) for i, v in enumerate(vals)]
    assert max(keys) == keys[4]                     # largest value wins
    assert keys[2] > keys[3]                        # tie: smaller row
    assert (keys[0] >> 5) == 0                      # exact zero: not a usable pivot
    k64 = [_pivot_key(np.float64(v), i) for i, v in enumerate([1e-300, 2.0, 2.0, 7.5])]
    assert max(k64) == k64[3] and k64[1] > k64[2]


def test_rhs_word_encoding(case):
    """the emitted RHS words decode (byte offset, FP64 high bits) to the
    tables' (slot, coefficient), as the kernel decodes them"""
    import struct
    _, tab, _, _, _, _ = case
    for row in tab.rhs_term:
        for (slot, c) in row:
            w = gpugen.rhs_word(slot, c)
            assert (w & 0xFFFF) == 8 * slot
            coef = struct.unpack("<d", struct.pack("<Q", (w & 0xFFFF0000) << 32))[0]
            assert coef == float(c)
    with pytest.raises(ValueError):
        gpugen.coef_hi16(1000003)   # not representable in 16 bits of the high word
