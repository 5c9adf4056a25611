"""Fixtures of the level-population routines, written by the reference's own
Fortran (READINPUT, FIND_CCOEFF, escape_probability, solvlevpop and
CALCULATE_DUST_TEMPERATURES compiled unchanged with amdflang into
oracle/_ref/ref_lvg_{1d,3d} by `make -C oracle ref`).

For each coolant (CII, CI, OI, CO) a set of seeded cases -- temperature,
dust temperature, density, collision partners, the point's populations and
chains of evaluation points with their populations along 12 rays -- including
the reference's special paths: a ray with no evaluation points, a level with
zero population (no line), a pair of levels exactly at their statistical
weights' ratio (|tpop| < 1e-50), inverted chains (masing, tau < -5), in the
3D build and in the PSEUDO_1D build (optical depth along ray 6 only).  Also
dust temperatures of seeded fields.

Writes tests/golden/lvg_fixtures.npz.  Run in the build container:
    python tests/golden/make_lvg_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "3d-pdr_amd"))
import oracle  # noqa: E402
from pdrchem import lamda  # noqa: E402

REF = "/root/reference"
NRAYS = 12
SPECIES_X = {"CII": 1.4e-4, "CI": 5e-5, "OI": 3e-4, "CO": 1e-4}


def lte(c, T, dens):
    w = c.weights * np.exp(-c.energies / lamda.KB / T)
    return dens * w / w.sum()


def make_cases(c, rng, ncases):
    cases = []
    for k in range(ncases):
        T = 10.0 ** rng.uniform(1.0, 3.3)
        rho = 10.0 ** rng.uniform(2.0, 6.0)
        x = SPECIES_X[c.name] * 10.0 ** rng.uniform(-2.0, 0.0)
        dens = [10.0 ** rng.uniform(-4, 0), 10.0 ** rng.uniform(-8, -3), 10.0 ** rng.uniform(-8, -3), 0.1,
                10.0 ** rng.uniform(-4, np.log10(0.5))]
        pop = lte(c, T * 10.0 ** rng.uniform(-0.3, 0.3), x * rho) * np.exp(0.2 * rng.normal(size=c.nlev))
        if k == 1:
            pop[2] = 0.0                                     # a level with no population
        if k == 2:
            pop[1] = pop[0] * c.weights[1] / c.weights[0]    # a pair at the weights' ratio
        rays = []
        for j in range(NRAYS):
            m = 0 if (k == 0 and j == 3) else int(rng.integers(1, 25))
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            pts = np.cumsum(np.vstack([np.zeros(3), d[None, :] * 10.0 ** rng.uniform(-3, -0.5, size=(m, 1))]), axis=0)
            Tc = T * 10.0 ** rng.uniform(-0.5, 0.3)
            pops = np.array([lte(c, Tc * 10.0 ** rng.uniform(-0.1, 0.1), x * rho * 10.0 ** rng.uniform(-0.5, 0.5))
                             for _ in range(m + 1)])
            if k == 3 and j in (5, 6):
                pops[:, 1:] *= 50.0                          # inverted upper levels: masing
            pops[0] = pop
            rays.append((pts, pops))
        cases.append(dict(T=T, Td=10.0 ** rng.uniform(1.0, 1.7), rho=rho, metallicity=1.0, v_turb=1.0e5, g2d=100.0,
                          dens=dens, xspec=x, pop=pop, rays=rays))
    return cases


def main():
    rng = np.random.default_rng(20251206)
    out = {}
    for cid, (name, fn, nlev, ntemp) in enumerate(lamda.COOLANTS, start=1):
        c = lamda.read_coolant(os.path.join(REF, fn), nlev, ntemp, name)
        for variant, p1 in (("3d", False), ("1d", True)):
            cases = make_cases(c, rng, 6)
            res = oracle.ref_lvg(os.path.join(REF, fn), nlev, ntemp, cid, cases, p1)
            key = f"{name}_{variant}"
            out[key + "_scalars"] = np.array([[cs["T"], cs["Td"], cs["rho"], cs["metallicity"], cs["v_turb"], cs["g2d"],
                                               cs["xspec"]] + list(cs["dens"]) for cs in cases])
            out[key + "_pop"] = np.array([cs["pop"] for cs in cases])
            # chains: per case and ray, the point count, then the points and populations
            out[key + "_epray"] = np.array([[len(p) - 1 for p, _ in cs["rays"]] for cs in cases])
            out[key + "_epoints"] = np.concatenate([np.concatenate([p for p, _ in cs["rays"]]) for cs in cases])
            out[key + "_chainpop"] = np.concatenate([np.concatenate([q for _, q in cs["rays"]]) for cs in cases])
            for f in ("C", "transition", "line", "tau", "beta", "solution"):
                out[key + "_" + f] = np.array([r[f] for r in res])
            out[key + "_cooling"] = np.array([r["cooling"] for r in res])
            print(key, "cooling", out[key + "_cooling"])
    # dust temperatures: attenuated field and per-ray surface fields
    uv = 10.0 ** rng.uniform(-3, 4, size=24)
    rs = np.zeros((24, NRAYS))
    for p in range(24):
        rs[p, rng.integers(0, NRAYS)] = 10.0 ** rng.uniform(-1, 5)
        if p % 3 == 0:
            rs[p] += 10.0 ** rng.uniform(-2, 2) / NRAYS
    uv[:2] = 0.0
    rs[0] = 0.0
    out["dust_uv"] = uv
    out["dust_rad_surface"] = rs
    out["dust_T"] = oracle.ref_dust_temperatures(uv, rs)
    np.savez_compressed(os.path.join(HERE, "lvg_fixtures.npz"), **out)


if __name__ == "__main__":
    main()
