"""The half-wavefront kernel (two systems per wavefront, launch shape
systems_per_block = 8) against the wavefront-per-system kernel: the same
abundances, statuses and integrator counters bit for bit (both run the same
arithmetic in the same order; only the lanes that carry it differ), on
random reduced-network points of uneven stiffness, an odd count (one half
runs out of points before the other), a single point, both precisions, the
two electron policies, and a step limit that makes points fail and retry
(the flattened output loop's retry/abort path)."""
import numpy as np
import pytest

from pdrchem import lib, workloads
from pdrchem.netgen import load_network

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    import torch
    net = load_network("reduced")
    nid = lib.network_id("reduced")
    n = 3001
    P = workloads.random_points(n, seed=11)
    dev = torch.device("cuda:0")
    rate, _ = lib.rates_device(nid, *[torch.tensor(v, device=dev) for v in (P.T, P.Td, P.G, P.Av, P.cols)])
    torch.cuda.synchronize()
    return net, nid, rate, torch.tensor(P.n_H, device=dev), workloads.initial_abundances(net, n)


def _solve(nid, rate, nH, x0, spb, precision, t_end=1e7):
    import torch
    dev = rate.device
    n = x0.shape[0]
    lib.clear_tuning()
    lib.set_tuned_config(nid, n, spb, 0, precision)
    ab = torch.tensor(x0, device=dev)
    stats = torch.zeros((n, 8), dtype=torch.int32, device=dev)
    status, _ = lib.solve_device(nid, ab, rate[:n].contiguous(), nH[:n].contiguous(), t_end=t_end,
                                 precision=precision, stats=stats)
    torch.cuda.synchronize()
    lib.clear_tuning()
    return ab.cpu().numpy(), status.cpu().numpy(), stats.cpu().numpy()


def _compare(got, ref, what):
    ab, st, sts = got
    ab0, st0, sts0 = ref
    assert np.array_equal(st, st0), what
    assert np.array_equal(sts, sts0), (what, np.argwhere(sts != sts0)[:5])
    same = (ab == ab0) | (np.isnan(ab) & np.isnan(ab0))
    assert np.all(same), (what, np.argwhere(~same)[:5])


@pytest.mark.parametrize("precision", [lib.MIXED, lib.FP64])
def test_pairs_equal_wavefronts(case, precision):
    net, nid, rate, nH, x0 = case
    for n in (3001, 1):
        print(f"pair vs wave: {n} points, precision {precision}", flush=True)
        ref = _solve(nid, rate, nH, x0[:n], 4, precision)
        got = _solve(nid, rate, nH, x0[:n], 8, precision)
        _compare(got, ref, (n, precision))
        assert np.all(ref[2][:, 0] > 0)


def test_pairs_equal_wavefronts_failures_and_electron_policy(case):
    """a step limit that stops points mid-run (CVode flag -1, retried up to
    five times), and the charge-sum electron policy"""
    net, nid, rate, nH, x0 = case
    n = 513
    old, pol = lib.mxstep(), lib.load().pdrchem_electron_policy()
    try:
        lib.set_mxstep(150)
        print("pair vs wave: mxstep 150", flush=True)
        ref = _solve(nid, rate, nH, x0[:n], 4, lib.MIXED)
        got = _solve(nid, rate, nH, x0[:n], 8, lib.MIXED)
        _compare(got, ref, "mxstep")
        assert np.any(ref[1] != 0) and np.any(ref[2][:, 7] == -1)
        lib.set_mxstep(old)
        lib.set_electron_policy(lib.XE_CHARGE_SUM if pol == lib.XE_LAST_RHS else lib.XE_LAST_RHS)
        print("pair vs wave: other electron policy", flush=True)
        ref = _solve(nid, rate, nH, x0[:n], 4, lib.MIXED, t_end=1e5)
        got = _solve(nid, rate, nH, x0[:n], 8, lib.MIXED, t_end=1e5)
        _compare(got, ref, "electron policy")
    finally:
        lib.set_mxstep(old)
        lib.set_electron_policy(pol)
