"""Reference solutions for the integrator checks, computed independently of
the oracle's BDF: scipy's Radau IIA (solve_ivp, rtol 1e-12; atol 1e-40 on the
reduced network, 1e-30 -- the reference's own -- on the full one, where 1e-40
makes Radau crawl through the trace species) with the oracle's exact
Jacobian, on the oracle's RHS (the RHS itself is pinned term by term against
the reference's odes_*.c in tests/test_netgen.py).

Writes tests/golden/truth_<net>.npz with the inputs (rates, n_H, x0, t_end)
and the Radau abundances at t_end for a handful of slab and random points.
Run in the build container:  python tests/golden/make_truth.py
"""
import os
import sys

import numpy as np
from scipy.integrate import solve_ivp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "3d-pdr_amd"))
import oracle  # noqa: E402
from pdrchem import workloads  # noqa: E402
from pdrchem.netgen import load_network  # noqa: E402

YR = 3.1556926e7


def main(net_name="reduced"):
    net = load_network(net_name)
    o = oracle.Oracle(net_name)
    slab = workloads.slab_1d(1000).subset(np.array([0, 333, 666, 999]))
    rnd = workloads.random_points(4, columns=(net_name == "full"))
    pts = [slab, rnd]
    atol = 1e-40 if net_name == "reduced" else 1e-30
    rates, nH, x0s, outs = [], [], [], []
    for P in pts:
        r = o.rates(P.T, P.Td, P.G, P.Av, P.cols, oracle.default_phys())
        x0 = workloads.initial_abundances(net, P.n)
        for k in range(P.n):
            f = lambda t, y, k=k: o.rhs(r[k], P.n_H[k], y)[0]
            jac = lambda t, y, k=k: o.jac(r[k], P.n_H[k], y)
            sol = solve_ivp(f, (0.0, 1.0e7 * YR), x0[k, :-1], method="Radau", rtol=1e-12, atol=atol,
                            jac=jac)
            assert sol.status == 0, sol.message
            rates.append(r[k])
            nH.append(P.n_H[k])
            x0s.append(x0[k])
            outs.append(sol.y[:, -1])
            print(net_name, len(outs), "nfev", sol.nfev, flush=True)
    np.savez_compressed(os.path.join(HERE, f"truth_{net_name}.npz"), rates=np.array(rates), n_H=np.array(nH),
                        x0=np.array(x0s), t_end_yr=1.0e7, radau=np.array(outs), atol=atol)


if __name__ == "__main__":
    main(*sys.argv[1:])
