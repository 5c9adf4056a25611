"""The allocation-free entry point (pdrchem_solve_device_noalloc) against
pdrchem_solve_device on the same inputs: two calls in a row, the second with
fewer points, each bitwise equal to the allocating entry point's result
(each point is integrated by its own wavefront or workgroup with a fixed
order of operations, whatever the workspace and block count)."""
import numpy as np
import pytest

import oracle
from pdrchem import lib, workloads
from pdrchem.netgen import load_network

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,npts,precisions", [("reduced", 600, (lib.MIXED, lib.FP64)),
                                                  ("full", 40, (lib.MIXED,))])
def test_noalloc_equals_allocating_entry(name, npts, precisions):
    import torch
    dev = torch.device("cuda:0")
    net = load_network(name)
    orc = oracle.Oracle(name)
    P = workloads.random_points(npts, seed=31, columns=(name != "reduced"))
    x0 = workloads.initial_abundances(net, npts)
    r = torch.tensor(orc.rates(P.T, P.Td, P.G, P.Av, P.cols), device=dev)
    nH = torch.tensor(P.n_H, device=dev)
    nid = lib.network_id(name)
    for prec in precisions:
        for n in (npts, npts // 3):  # second call smaller: the scratch keeps the first call's contents
            print(f"case: noalloc {name} {n} points precision={prec}", flush=True)
            a_ref = torch.tensor(x0[:n], device=dev)
            st_ref = torch.zeros((n, 8), dtype=torch.int32, device=dev)
            s_ref, _ = lib.solve_device(nid, a_ref, r[:n].contiguous(), nH[:n].contiguous(), stats=st_ref,
                                        precision=prec)
            a = torch.tensor(x0[:n], device=dev)
            st = torch.zeros((n, 8), dtype=torch.int32, device=dev)
            s, _ = lib.solve_device_noalloc(nid, a, r[:n].contiguous(), nH[:n].contiguous(), stats=st,
                                            precision=prec)
            torch.cuda.synchronize()
            assert int(s_ref.abs().sum().item()) == 0
            assert torch.equal(s, s_ref)
            assert torch.equal(a, a_ref)
            assert torch.equal(st, st_ref)


def test_noalloc_rejects_bad_arguments():
    import torch
    dev = torch.device("cuda:0")
    a = torch.zeros((4, 33), dtype=torch.float64, device=dev)
    r = torch.zeros((4, 329), dtype=torch.float64, device=dev)
    d = torch.ones(4, dtype=torch.float64, device=dev)
    with pytest.raises(lib.PdrChemError):
        lib.solve_device_noalloc(7, a, r, d)
    with pytest.raises(lib.PdrChemError):
        lib.solve_device_noalloc(0, a, r, d, nelect=99)
