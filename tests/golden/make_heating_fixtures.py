"""Write golden heating vectors from the reference's own CALC_HEATING.

Runs the amdflang-built reference CALC_HEATING (sub_calculate_heating.F90 with
the reference's modules, driven by oracle/ref_heating_driver.F90; built by
`make -C oracle ref` as oracle/_ref/ref_heating_<network>) on seeded points
spanning cold molecular to hot ionised gas (n_H 1e1..1e6 cm^-3, T 10..1e4 K,
T_dust 5..150 K, UV 1e-3..1e5 Draine, abundances log-uniform 1e-12..1e-1),
with the rate coefficients the oracle computes for those conditions.
Output: tests/golden/heating_<network>.npz with the inputs and the 12 heating
terms per point.

Run in the build container (the reference tree and amdflang are needed):
    python tests/golden/make_heating_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "3d-pdr_amd")]
import oracle  # noqa: E402

NPTS = {"reduced": 48, "full": 24}


def inputs(name, seed):
    rng = np.random.default_rng(seed)
    o = oracle.Oracle(name)
    n = 10 ** rng.uniform(1, 6, NPTS[name])
    T = 10 ** rng.uniform(1, 4, NPTS[name])
    Td = rng.uniform(5, 150, NPTS[name])
    uv = 10 ** rng.uniform(-3, 5, NPTS[name])
    Av = rng.uniform(0, 20, (NPTS[name], 12))
    G = np.outer(uv, np.ones(12)) / 12.0
    cols = 10 ** rng.uniform(10, 22, (NPTS[name], 12, 5))
    rate = o.rates(T, Td, G, Av, cols)
    ab = 10 ** rng.uniform(-12, -1, (NPTS[name], o.nspec))
    return n, T, Td, uv, ab, rate


def main():
    for i, name in enumerate(("reduced", "full")):
        n, T, Td, uv, ab, rate = inputs(name, 20251207 + i)
        heat = oracle.ref_heating(name, n, T, Td, uv, ab, rate)
        np.savez_compressed(os.path.join(HERE, f"heating_{name}.npz"), n=n, T=T, Td=Td, uv=uv, ab=ab, rate=rate,
                            heat=heat)
        print(name, heat.shape)


if __name__ == "__main__":
    main()
