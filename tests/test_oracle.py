"""Pins of the CPU oracle (oracle/), independent of the HIP path.

* rates: against golden vectors written by the reference's own Fortran rate
  code (tests/golden/rates_*.npz, see make_rate_fixtures.py) and, when the
  amdflang-built reference is present (oracle/_ref), against fresh random
  inputs; closed forms for simple reaction types.
* RHS / Jacobian: element conservation of ydot, charge sum, exactness of the
  Jacobian (central differences are exact for the quadratic RHS).
* integrator: closed-form linear decay, the Robertson problem's published
  solution, and independent Radau solutions (tests/golden/truth_*.npz).
"""
import os

import numpy as np
import pytest

import oracle
from pdrchem import netgen, workloads
from pdrchem.netgen import load_network

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _phys_from(a):
    return oracle.Phys(*[float(v) for v in a])


@pytest.mark.parametrize("net", ["reduced", "full", "mynetwork"])
def test_rates_match_reference_golden(net):
    d = np.load(os.path.join(GOLD, f"rates_{net}.npz"))
    o = oracle.Oracle(net)
    r = o.rates(d["T"], d["Td"], d["G"], d["Av"], d["cols"], _phys_from(d["phys"]))
    ref = d["rates"]
    both_zero = (r == 0) & (ref == 0)
    rel = np.where(both_zero, 0.0, np.abs(r - ref) / np.maximum(np.abs(ref), 1e-300))
    assert rel.max() < 1e-11, (rel.max(), np.unravel_index(rel.argmax(), rel.shape))
    # the golden set exercises every reaction type the network has
    types = {r_.rtype for r_ in load_network(net).reactions}
    assert types  # sanity
    assert np.count_nonzero(ref) > 0.5 * ref.size


@pytest.mark.parametrize("net", ["reduced", "full"])
@pytest.mark.skipif(not oracle.have_ref_rates("reduced") or not os.path.isdir("/root/reference"),
                    reason="reference rate code not built here")
def test_rates_match_reference_random(net):
    rng = np.random.default_rng(7)
    n = 64
    T = 10 ** rng.uniform(0.5, 4.5, n)
    Td = rng.uniform(5, 80, n)
    G = 10 ** rng.uniform(-1, 5, (n, 3))
    Av = rng.uniform(0, 30, (n, 3))
    cols = 10 ** rng.uniform(10, 24, (n, 3, 5))
    o = oracle.Oracle(net)
    a = o.rates(T, Td, G, Av, cols)
    b = oracle.ref_rates(net, T, Td, G, Av, cols)
    rel = np.where((a == 0) & (b == 0), 0.0, np.abs(a - b) / np.maximum(np.abs(b), 1e-300))
    assert rel.max() < 1e-11


def test_rate_closed_forms():
    net = load_network("reduced")
    o = oracle.Oracle("reduced")
    ph = oracle.default_phys()
    T = np.array([100.0])
    r = o.rates(T, 20.0, [[0.0]], [[0.0]], None, ph)[0]
    for rx in net.reactions:
        i = rx.index
        if rx.rtype == netgen.T_CRP:
            assert r[i] == pytest.approx(rx.alpha * ph.zeta if rx.alpha * ph.zeta > 1e-99 else 0.0, rel=1e-15)
        if rx.rtype in (netgen.T_PHOTO, netgen.T_PHOTO_CO, netgen.T_PHOTO_CI, netgen.T_PHOTO_H2):
            assert r[i] == 0.0  # no field, no photo-reactions
        if rx.rtype == netgen.T_THERMAL and rx.duplicate == 0 and not (rx.gamma < -200 and 100 < rx.rtmin):
            v = rx.alpha * (100.0 / 300.0) ** rx.beta * np.exp(-rx.gamma / 100.0)
            v = min(v, 1.0)
            v = 0.0 if v < 1e-99 else v
            assert r[i] == pytest.approx(v, rel=1e-13)
    # unshielded generic photo-rate: alpha*G*exp(-gamma*Av)/2
    r2 = o.rates(T, 20.0, [[10.0]], [[1.5]], None, ph)[0]
    for rx in net.reactions:
        if rx.rtype == netgen.T_PHOTO and rx.duplicate == 0:
            assert r2[rx.index] == pytest.approx(rx.alpha * 10.0 * np.exp(-rx.gamma * 1.5) / 2.0, rel=1e-14)


def _composition_matrix(net):
    import test_netgen
    els = sorted({e for s in net.species[:-1] for e in (test_netgen.composition(s.name) or {})})
    E = np.zeros((len(els), net.neq))
    ok = np.ones(net.neq, bool)
    for i, s in enumerate(net.species[:-1]):
        c = test_netgen.composition(s.name) if not s.name.startswith("PAH") else None
        if c is None:
            ok[i] = False
            continue
        for e, k in c.items():
            E[els.index(e), i] = k
    return E, ok


def test_rhs_conserves_elements_and_jacobian_exact():
    net = load_network("reduced")
    o = oracle.Oracle("reduced")
    rng = np.random.default_rng(3)
    r = o.rates([80.0], 20.0, [[30.0]], [[2.0]])[0]
    E, ok = _composition_matrix(net)
    assert ok.all()
    q = np.array(net.charges(), dtype=float)
    for _ in range(5):
        x = 10 ** rng.uniform(-20, 0, net.neq)
        yd, xe = o.rhs(r, 1e3, x)
        assert xe == pytest.approx(float(q @ x), rel=1e-14)
        scale = np.abs(E) @ np.abs(yd) + 1e-300
        assert np.all(np.abs(E @ yd) <= 1e-12 * scale + 1e-300)
        J = o.jac(r, 1e3, x)
        Jfd = np.zeros_like(J)
        h = 1e-3
        for j in range(net.neq):
            xp, xm = x.copy(), x.copy()
            xp[j] += h
            xm[j] -= h
            Jfd[:, j] = (o.rhs(r, 1e3, xp)[0] - o.rhs(r, 1e3, xm)[0]) / (2 * h)
        assert np.max(np.abs(J - Jfd)) <= 1e-9 * np.max(np.abs(J))


def test_integrator_linear_decay():
    for mode in (oracle.MODE_REFERENCE, oracle.MODE_GPU, oracle.MODE_GPU_MIXED):
        t = np.array([0.5, 1.0, 2.0, 4.0])
        flag, y, s = oracle.integrate_generic(lambda t, y: -2.0 * y, lambda t, y: np.array([[-2.0]]),
                                              [1.0], t, rtol=1e-8, atol=1e-30, mode=mode)
        assert flag == 0
        assert np.allclose(y[:, 0], np.exp(-2.0 * t), rtol=1e-6)


def test_integrator_robertson():
    """Robertson's problem; values at t=40 from the classical literature
    (Hairer & Wanner, Solving ODEs II): y1 = 0.7158270687, y3 = 0.2841637277."""
    def f(t, y):
        return np.array([-0.04 * y[0] + 1e4 * y[1] * y[2],
                         0.04 * y[0] - 1e4 * y[1] * y[2] - 3e7 * y[1] ** 2, 3e7 * y[1] ** 2])

    def J(t, y):
        return np.array([[-0.04, 1e4 * y[2], 1e4 * y[1]], [0.04, -1e4 * y[2] - 6e7 * y[1], -1e4 * y[1]],
                         [0, 6e7 * y[1], 0]])

    for mode in (oracle.MODE_REFERENCE, oracle.MODE_GPU):
        flag, y, s = oracle.integrate_generic(f, J, [1.0, 0.0, 0.0], [40.0], rtol=1e-10, atol=1e-14, mode=mode)
        assert flag == 0
        assert y[0, 0] == pytest.approx(0.7158270687, rel=1e-7)
        assert y[0, 2] == pytest.approx(0.2841637277, rel=1e-7)
        assert np.sum(y[0]) == pytest.approx(1.0, rel=1e-9)


@pytest.mark.parametrize("mode", ["MODE_GPU", "MODE_GPU_MIXED"])
def test_integrator_vs_radau(mode):
    path = os.path.join(GOLD, "truth_reduced.npz")
    if not os.path.exists(path):
        pytest.skip("truth fixture not generated")
    d = np.load(path)
    o = oracle.Oracle("reduced")
    st, out, stats = o.calculate_abundances(d["x0"].copy(), d["rates"], d["n_H"], None,
                                            mode=getattr(oracle, mode), nthreads=4)
    assert st == 0
    truth = d["radau"]
    m = truth > 1e-20
    rel = np.abs(out[:, :-1] - truth)[m] / truth[m]
    assert rel.max() < 1e-5, rel.max()


@pytest.mark.parametrize("mode", ["MODE_GPU_WG", "MODE_GPU_WG_MIXED", "MODE_GPU"])
def test_integrator_vs_radau_full(mode):
    """the full network's integrator configurations against independent
    Radau IIA solutions (rtol 1e-12, atol 1e-30; make_truth.py full) on four
    slab and four random points with columns: all within 1e-6"""
    d = np.load(os.path.join(GOLD, "truth_full.npz"))
    o = oracle.Oracle("full")
    st, out, _ = o.calculate_abundances(d["x0"].copy(), d["rates"], d["n_H"], None, mode=getattr(oracle, mode),
                                        nthreads=4)
    assert st == 0
    truth = d["radau"]
    m = truth > 1e-20
    rel = np.abs(out[:, :-1] - truth)[m] / truth[m]
    assert rel.max() < 5e-6, rel.max()


def test_calculate_abundances_semantics():
    """clip below atol, electron slot from the charge sum, status 0."""
    net = load_network("reduced")
    o = oracle.Oracle("reduced")
    P = workloads.slab_1d(1000).subset(np.array([10, 900]))
    r = o.rates(P.T, P.Td, P.G, P.Av, P.cols)
    x0 = workloads.initial_abundances(net, P.n)
    st, out, stats = o.calculate_abundances(x0, r, P.n_H, P.T, mode=oracle.MODE_GPU, nthreads=2)
    assert st == 0
    assert np.all((out[:, :-1] == 0) | (out[:, :-1] >= 1e-30))
    q = np.array(net.charges(), dtype=float)
    assert np.allclose(out[:, -1], out[:, :-1] @ q, rtol=1e-6)
    assert np.all(stats["nst"] > 0)


# ---------------------------------------------------------------------------
# failure policy (calculate_abundances.c:216-317) and the electron slot
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def slab8():
    net = load_network("reduced")
    o = oracle.Oracle("reduced")
    P = workloads.slab_1d(1000).subset(np.arange(0, 1000, 125))
    r = o.rates(P.T, P.Td, P.G, P.Av, P.cols)
    return net, o, P, r, workloads.initial_abundances(net, P.n)


def test_retries_continue_the_same_integration(slab8):
    """A CVode call that runs out of steps (-1) is called again with the same
    budget: a point that needs at most four such retries ends bitwise where
    an unlimited run ends; the fifth failure ends the point with -1 and
    writes back the state reached, short of t_end."""
    net, o, P, r, x0 = slab8
    st0, ref0, s0 = o.calculate_abundances(x0.copy(), r, P.n_H, P.T, mode=oracle.MODE_GPU, nthreads=8)
    assert st0 == 0 and np.all(s0["nfails"] == 0)
    st, ab, s = o.calculate_abundances(x0.copy(), r, P.n_H, P.T, mode=oracle.MODE_GPU, nthreads=8, mxstep=300)
    assert st == -1
    rec = (s["flag"] == 0) & (s["nfails"] > 0)
    exh = s["flag"] == -1
    assert rec.any() and exh.any()
    assert np.all(s["nfails"][exh] == 5) and np.all(s["nfails"][~exh] < 5)
    for k in np.where(~exh)[0]:
        assert np.array_equal(ab[k], ref0[k])
    for k in np.where(exh)[0]:
        assert s["nst"][k] < s0["nst"][k]
        assert not np.array_equal(ab[k], ref0[k])
        assert np.all(np.isfinite(ab[k]))


def test_nan_rate_fails_its_point_only(slab8):
    net, o, P, r, x0 = slab8
    rn = r.copy()
    rn[3, 10] = np.nan
    _, ref0, _ = o.calculate_abundances(x0.copy(), r, P.n_H, P.T, mode=oracle.MODE_GPU, nthreads=8, mxstep=5000)
    st, ab, s = o.calculate_abundances(x0.copy(), rn, P.n_H, P.T, mode=oracle.MODE_GPU, nthreads=8, mxstep=5000)
    assert st == -4 and s["flag"][3] == -4 and s["nst"][3] == 0 and s["nfails"][3] == 5
    others = np.arange(P.n) != 3
    assert np.array_equal(ab[others], ref0[others])


def test_electron_policy(slab8):
    """policy 1 writes the charge sum of the returned abundances; policy 0
    (the reference's) the last RHS evaluation's x_e, which differs from it
    by the last step's overshoot past t_end."""
    net, o, P, r, x0 = slab8
    q = np.array(net.charges(), dtype=np.float64)
    _, a0, _ = o.calculate_abundances(x0.copy(), r, P.n_H, P.T, mode=oracle.MODE_GPU, nthreads=8)
    _, a1, _ = o.calculate_abundances(x0.copy(), r, P.n_H, P.T, mode=oracle.MODE_GPU, nthreads=8, xe_policy=1)
    assert np.array_equal(a0[:, :-1], a1[:, :-1])
    assert np.allclose(a1[:, -1], a1[:, :-1] @ q, rtol=1e-14, atol=0)
    assert not np.array_equal(a0[:, -1], a1[:, -1])
    assert np.all(np.abs(a0[:, -1] - a1[:, -1]) <= 1e-2 * a1[:, -1])


def test_hybrid_head_pivot_rule():
    """The bordered factorisation eliminates its head without pivoting: a
    zero or non-finite head pivot is a setup failure (the device kernel's
    rule), in FP64 and in FP32 (where a pivot below the smallest normal
    float also fails); a regular matrix factorises."""
    L = oracle.lib()
    nb = 4
    order = np.array([0, 1, 2, 3], dtype=np.int32)
    ip = order.ctypes.data_as(oracle.C.POINTER(oracle.C.c_int))
    base = np.array([[4.0, 1.0, 0.0, 0.5], [1.0, 3.0, 1.0, 0.0], [0.0, 1.0, 2.0, 1.0], [0.5, 0.0, 1.0, 5.0]])

    def status(a, f32=0):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel().copy()
        return L.pdr_test_hybrid_factor(nb, a, ip, 2, f32)

    assert status(base) == 0 and status(base, 1) == 0
    z = base.copy()
    z[0, 0] = 0.0
    assert status(z) == 1 and status(z, 1) == 1
    z = base.copy()
    z[1, 1] = 0.25  # second head pivot 0.25 - 1/4 * 1 = 0 after the first elimination
    assert status(z) == 1
    z = base.copy()
    z[0, 0] = np.nan
    assert status(z) == 1
    z = base.copy()
    z[0, 0] = np.inf
    assert status(z) == 1
    z = base.copy()
    z[0, 0] = 1e-40  # a normal double, a denormal float
    assert status(z) == 0 and status(z, 1) == 1
