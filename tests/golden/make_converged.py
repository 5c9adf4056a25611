"""Converged reference solutions for GPU parity on the seeded random-point
workload (reduced network): the CPU oracle at the matched tolerances
(rtol 1e-8, atol 1e-30, GPU-equivalent integrator configuration) and the
same oracle at rtol 1e-11 / atol 1e-33 as the converged solution.

Writes tests/golden/converged_reduced.npz.  Run in the build container:
    python tests/golden/make_converged.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "3d-pdr_amd"))
import oracle  # noqa: E402
from pdrchem import workloads  # noqa: E402
from pdrchem.netgen import load_network  # noqa: E402


def main(n=96):
    net = load_network("reduced")
    o = oracle.Oracle("reduced")
    P = workloads.random_points(n)
    x0 = workloads.initial_abundances(net, P.n)
    r = o.rates(P.T, P.Td, P.G, P.Av, P.cols)
    st, ref, stats = o.calculate_abundances(x0.copy(), r, P.n_H, P.T, mode=oracle.MODE_GPU, nthreads=8)
    assert st == 0
    st, conv, cst = o.calculate_abundances(x0.copy(), r, P.n_H, P.T, rtol=1e-11, atol=1e-33,
                                           mode=oracle.MODE_GPU, nthreads=8)
    assert st == 0
    np.savez_compressed(os.path.join(HERE, "converged_reduced.npz"), rates=r, n_H=P.n_H, T=P.T, x0=x0,
                        oracle=ref, converged=conv, oracle_nst=stats["nst"], converged_nst=cst["nst"])
    print("done", stats["nst"].mean(), cst["nst"].mean())


if __name__ == "__main__":
    main()
