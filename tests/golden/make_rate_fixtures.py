"""Write golden rate-coefficient vectors from the reference's own rate code.

Runs the amdflang-built reference CALCULATE_REACTION_RATES (oracle/_ref,
built by `make -C oracle ref` from the reference sources) on seeded inputs
that cover every branch the networks use: zero and strong fields, zero and
large extinction, zero and large columns (H2/CO/C self-shielding), and
temperatures on both sides of every duplicate window and of the gamma<-200
cutoff.  Output: tests/golden/rates_<net>.npz (inputs + rates).

Run in the build container (the reference tree and amdflang are needed):
    python tests/golden/make_rate_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

NPTS = {"reduced": 48, "full": 16, "mynetwork": 24}


def inputs(n, seed):
    rng = np.random.default_rng(seed)
    T = 10 ** rng.uniform(0.5, 4.6, n)
    # hit duplicate windows and cutoffs exactly at some points
    T[:6] = [10.0, 300.0, 1000.0, 2000.0, 41000.0, 7.0]
    Td = rng.uniform(5.0, 90.0, n)
    nrays = 2
    G = 10 ** rng.uniform(-1, 5, (n, nrays))
    G[rng.random((n, nrays)) < 0.2] = 0.0
    Av = rng.uniform(0.0, 25.0, (n, nrays))
    Av[rng.random((n, nrays)) < 0.2] = 0.0
    cols = 10 ** rng.uniform(10, 23.5, (n, nrays, 5))
    cols[rng.random((n, nrays, 5)) < 0.25] = 0.0
    return T, Td, G, Av, cols


def main():
    for net, n in NPTS.items():
        if not oracle.have_ref_rates(net):
            print(f"skip {net}: oracle/_ref not built")
            continue
        T, Td, G, Av, cols = inputs(n, 1234 + n)
        ph = oracle.default_phys()
        rates = oracle.ref_rates(net, T, Td, G, Av, cols, ph)
        np.savez_compressed(os.path.join(HERE, f"rates_{net}.npz"), T=T, Td=Td, G=G, Av=Av, cols=cols,
                            rates=rates,
                            phys=np.array([ph.zeta, ph.v_turb, ph.omega, ph.grain_radius, ph.metallicity,
                                           ph.g2d]))
        print(net, rates.shape)


if __name__ == "__main__":
    main()
