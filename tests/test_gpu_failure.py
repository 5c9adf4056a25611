"""Failure policy of calculate_abundances_ on the device (calculate_abundances.c:
216-317): CVode calls that run out of steps (-1) are retried with the same
step budget up to five failures; other flags abort the point; the state
reached is written back; failed points are logged to cvode.log.  The flags
are forced from the host (a small mxstep, a NaN rate coefficient), checked
against the oracle run with the same settings.  Also the request loop's
batch cap lowered below a queued request.
"""
import os

import numpy as np
import pytest

import oracle
from parity import composition_matrix
from pdrchem import lib, workloads
from pdrchem.netgen import load_network

pytestmark = pytest.mark.gpu


def max_rel(a, ref, floor=1e-20):
    m = ref > floor
    return float(np.max(np.abs(a[m] - ref[m]) / ref[m])) if np.any(m) else 0.0


@pytest.fixture(scope="module")
def case():
    net = load_network("reduced")
    orc = oracle.Oracle("reduced")
    P = workloads.slab_1d(1000).subset(np.arange(0, 1000, 125))  # 8 points, 646..2562 steps each
    r = orc.rates(P.T, P.Td, P.G, P.Av, P.cols)
    x0 = workloads.initial_abundances(net, P.n)
    return net, orc, P, r, x0


@pytest.fixture(autouse=True)
def _restore_options():
    yield
    lib.set_mxstep(10_000_000)
    lib.set_electron_policy(lib.XE_LAST_RHS)


def _solve(x0, r, P, precision, stats=True):
    import torch
    dev = torch.device("cuda:0")
    ab = torch.tensor(x0, device=dev)
    st = torch.zeros((P.n, 8), dtype=torch.int32, device=dev) if stats else None
    status, _ = lib.solve_device(lib.network_id("reduced"), ab, torch.tensor(r, device=dev),
                                 torch.tensor(P.n_H, device=dev), stats=st, precision=precision)
    torch.cuda.synchronize()
    return status.cpu().numpy(), ab.cpu().numpy(), (st.cpu().numpy() if stats else None)


def test_mxstep_retry_and_exhaustion(case):
    """mxstep 300: the oracle retries some points (1..4 failed CVode calls)
    and finishes them with results bitwise equal to an unlimited run, and
    exhausts the others (status -1 after the fifth failure, partial state
    written back).  The device must give the same statuses, bitwise its own
    unlimited result on the recovered points, and on the exhausted ones a
    partial state that is a state of the same integration: every element
    with a total above 1e-10 conserved to 1e-6, short of the final state.  (The
    partial state itself is ill-conditioned -- it is wherever in time the
    fifth failure stopped the point, and device and oracle step sequences
    drift apart at the 1e-7 level of the step-size roots.)"""
    net, orc, P, r, x0 = case
    E, ok = composition_matrix(net)
    assert ok.all()
    for prec, mode in ((lib.FP64, oracle.MODE_GPU), (lib.MIXED, oracle.MODE_GPU_MIXED)):
        st0, ref0, _ = orc.calculate_abundances(x0.copy(), r, P.n_H, P.T, mode=mode, nthreads=8)
        st1, ref1, s1 = orc.calculate_abundances(x0.copy(), r, P.n_H, P.T, mode=mode, nthreads=8, mxstep=300)
        assert st0 == 0 and st1 == -1
        # points whose outcome does not change with mxstep 285 or 315: a
        # point that needs just about 300 steps in one call may go either way
        # on integrators whose rounding differs
        flags = [orc.calculate_abundances(x0.copy(), r, P.n_H, P.T, mode=mode, nthreads=8, mxstep=m)[2]["flag"]
                 for m in (285, 315)]
        stable = (flags[0] == s1["flag"]) & (flags[1] == s1["flag"])
        rec = stable & (s1["flag"] == 0) & (s1["nfails"] > 0)
        exh = stable & (s1["flag"] != 0)
        assert rec.any() and exh.any(), (s1["nfails"], s1["flag"], stable)
        for k in np.where(s1["flag"] == 0)[0]:
            assert np.array_equal(ref1[k], ref0[k])
        print(f"case: slab 8 points mxstep=300 precision={prec}", flush=True)
        lib.set_mxstep(10_000_000)
        g0, a0, _ = _solve(x0, r, P, prec)
        lib.set_mxstep(300)
        g1, a1, gs = _solve(x0, r, P, prec)
        print("  oracle nfails", s1["nfails"].tolist(), "flags", s1["flag"].tolist())
        print("  device status", g1.tolist(), "nst", gs[:, 0].tolist())
        assert np.all(g0 == 0)
        assert np.array_equal(g1[stable], s1["flag"][stable]), (g1, s1["flag"], stable)
        assert np.all((g1 == 0) | (g1 == -1))
        for k in np.where(g1 == 0)[0]:
            assert np.array_equal(a1[k], a0[k])  # retried calls continue the same integration
        for k in np.where(exh)[0]:
            y, y0 = a1[k, :-1], x0[k, :-1]
            tot = E @ y0
            el = (np.abs(E @ y - tot) / tot)[tot > 1e-10]
            print(f"  exhausted point {k}: element drift {el.max():.1e}, "
                  f"distance from the final state {max_rel(y, ref0[k, :-1]):.1e}")
            assert np.all(np.isfinite(y)) and el.max() <= 1e-6
            assert max_rel(y, ref0[k, :-1]) > 1e-3  # it did stop short of t_end


def test_nan_rate_fails_only_its_point(case):
    """A NaN rate coefficient: the point fails with -4 (convergence failures
    on the first step, five times; what it writes back is the failed first
    step's state, NaN where the reaction reaches), and the other points of
    the batch are bitwise unaffected.  mxstep is kept small so the failing
    point's work is bounded whatever its flag."""
    net, orc, P, r, x0 = case
    rn = r.copy()
    rn[3, 10] = np.nan
    st, ref, s = orc.calculate_abundances(x0.copy(), rn, P.n_H, P.T, mode=oracle.MODE_GPU, nthreads=8,
                                          mxstep=5000)
    assert st == -4 and s["flag"][3] == -4 and s["nst"][3] == 0
    for prec in (lib.FP64, lib.MIXED):
        print(f"case: slab 8 points, NaN rate at point 3, mxstep 5000, precision={prec}", flush=True)
        lib.set_mxstep(5000)
        g0, a0, _ = _solve(x0, r, P, prec)
        g1, a1, gs = _solve(x0, rn, P, prec)
        print("  device status", g1.tolist(), "nst", gs[:, 0].tolist())
        assert g1[3] == -4 and gs[3, 0] == 0
        others = np.arange(P.n) != 3
        assert np.all(g1[others] == 0)
        assert np.array_equal(a1[others], a0[others])


def test_calculate_abundances_status_and_log(case, tmp_path, monkeypatch):
    """The drop-in entry point returns the failed point's flag and appends the
    reference's fifth-failure block for every exhausted point to ./cvode.log."""
    net, orc, P, r, x0 = case
    monkeypatch.chdir(tmp_path)
    lib.set_globals(rtol=1e-8, atol=1e-30, start_time=0.0, end_time=1e7, nelect=33)
    lib.set_mxstep(300)
    ab = x0.copy()
    print("case: calculate_abundances_ with mxstep 300", flush=True)
    st = lib.calculate_abundances(ab, np.ascontiguousarray(r), P.n_H.copy(), P.T.copy(), precision=lib.FP64)
    assert st == -1
    # the points that failed, from the device's own statuses
    g1, _, _ = _solve(x0, r, P, lib.FP64, stats=False)
    log = (tmp_path / "cvode.log").read_text()
    exh = [int(k) for k in np.where(g1 != 0)[0]]
    assert exh and log.count("ERROR! Fifth failure - aborting integration") == len(exh)
    for k in exh:
        assert f"Particle {k}\n" in log
    assert "flag   = -1" in log
    lib.set_mxstep(10_000_000)
    ab = x0.copy()
    assert lib.calculate_abundances(ab, np.ascontiguousarray(r), P.n_H.copy(), P.T.copy(), precision=lib.MIXED) == 0
    assert log == (tmp_path / "cvode.log").read_text()  # nothing appended on success


def test_server_cap_lowered_below_queued_request(case):
    """A request queued before the batch cap is lowered below its size still
    runs (alone in its batch) and progress(wait) returns."""
    net, orc, P, r, x0 = case
    S = lib.Server(nslots=1, precision=lib.MIXED)
    try:
        a1 = np.ascontiguousarray(np.repeat(x0, 4, axis=0))
        r1 = np.ascontiguousarray(np.repeat(r, 4, axis=0))
        d1 = np.ascontiguousarray(np.repeat(P.n_H, 4))
        a2, r2, d2 = a1.copy(), r1.copy(), d1.copy()
        t1 = S.submit(0, a1, r1, d1, 1e7)   # occupies the only slot
        t2 = S.submit(0, a2, r2, d2, 1e7)   # waits in the queue
        S.set_max_batch(8)                  # below the queued request's 32 points
        print("case: server, 1 slot, cap 8 < queued 32 points", flush=True)
        assert S.progress(wait=True) == 0
        done = {t: st for t, st, _ in S.poll()}
        assert done == {t1: 0, t2: 0}
        assert np.array_equal(a1, a2)
    finally:
        S.close()
