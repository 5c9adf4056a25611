"""Autotuner of the solver's launch shape (pdrchem_autotune): every shape it
may choose gives the same results bitwise (each point is integrated by its
own wavefront / workgroup with a fixed order of operations, whatever the
block shape), the tuner leaves the batch unchanged, records its choice for
the batch's size bucket only, and later launches in that bucket use it."""
import numpy as np
import pytest
import torch

from pdrchem import lib, workloads
from pdrchem.netgen import load_network

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def batch(name, n, seed):
    net = load_network(name)
    nid = lib.network_id(name)
    P = workloads.random_points(n, seed=seed, columns=(name != "reduced"))
    rate, neg = lib.rates_device(nid, *[torch.tensor(v, device=DEV) for v in (P.T, P.Td, P.G, P.Av, P.cols)])
    x0 = torch.tensor(workloads.initial_abundances(net, n), device=DEV)
    return nid, x0, rate, torch.tensor(P.n_H, device=DEV)


def solve(nid, x0, rate, nH, t_end):
    ab = x0.clone()
    status, _ = lib.solve_device(nid, ab, rate, nH, t_end=t_end, precision=lib.MIXED)
    torch.cuda.synchronize()
    assert int(status.abs().sum().item()) == 0
    return ab.cpu().numpy()


@pytest.mark.parametrize("name,n,shapes", [
    ("reduced", 300, [(1, 0), (2, 0), (4, 0), (4, 2), (1, 3)]),
    ("full", 6, [(1, 1), (1, 2)]),
])
def test_every_shape_same_result(name, n, shapes):
    print(f"autotune shapes {name} n={n}", flush=True)
    nid, x0, rate, nH = batch(name, n, seed=99)
    t_end = 1e5 if name == "full" else 1e7
    lib.clear_tuning()
    ref = solve(nid, x0, rate, nH, t_end)
    for spb, bpc in shapes:
        lib.set_tuned_config(nid, n, spb, bpc)
        assert lib.tuned_config(nid, n) == (spb, bpc)
        got = solve(nid, x0, rate, nH, t_end)
        assert np.array_equal(got, ref), (spb, bpc)
    lib.set_tuned_config(nid, n, 0, 0)
    assert lib.tuned_config(nid, n) is None


@pytest.mark.parametrize("name,n", [("reduced", 2000), ("full", 12)])
def test_autotune_records_choice_and_keeps_batch(name, n):
    print(f"autotune {name} n={n}", flush=True)
    nid, x0, rate, nH = batch(name, n, seed=5)
    t_end = 1e5 if name == "full" else 1e7
    lib.clear_tuning()
    ref = solve(nid, x0, rate, nH, t_end)
    ab = x0.clone()
    r = lib.autotune(nid, ab, rate, nH, t_end=t_end)
    torch.cuda.synchronize()
    assert torch.equal(ab, x0)                       # the batch is restored
    cands = {(4, 0), (2, 0), (1, 0), (4, 2)} if name == "reduced" else {(1, 0), (1, 1)}
    assert (r["systems_per_block"], r["blocks_per_cu"]) in cands and r["ms"] > 0
    cfg = lib.tuned_config(nid, n)
    assert cfg == (r["systems_per_block"], r["blocks_per_cu"]) or (cfg is None and r["blocks_per_cu"] == 0
                                                                   and r["systems_per_block"] == 0)
    assert lib.tuned_config(nid, max(n // 16, 1)) is None          # smaller buckets keep the default
    assert lib.tuned_config(nid, n * 16) == lib.tuned_config(nid, n)  # larger ones follow the tuned bucket
    got = solve(nid, x0, rate, nH, t_end)            # launched with the tuned shape
    assert np.array_equal(got, ref)
    lib.clear_tuning()
