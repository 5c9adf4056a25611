"""Pins of the oracle's level-population and cooling routines (oracle/thermal.c)
and of the coolant reader (pdrchem/lamda.py):

* against the reference's own Fortran (READINPUT, FIND_CCOEFF,
  escape_probability, solvlevpop, CALCULATE_DUST_TEMPERATURES built unchanged
  by oracle/Makefile) through the committed fixtures
  (tests/golden/lvg_fixtures.npz, make_lvg_fixtures.py), 3D and PSEUDO_1D,
  and live when the build is present;
* closed forms: LTE populations sum to the density with Boltzmann ratios;
  at densities far above the critical density the statistical equilibrium
  is the LTE distribution; optically thin lines (no chain) have beta = 1.
"""
import os

import numpy as np
import pytest

import oracle
from pdrchem import lamda

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = "/root/reference"


@pytest.fixture(scope="module")
def tables():
    return {c.name: oracle.CoolantTables(c) for c in lamda.load_coolants()}


@pytest.fixture(scope="module")
def fx():
    return np.load(os.path.join(GOLD, "lvg_fixtures.npz"))


def fixture_cases(fx, key):
    sc, pop, ep, epts, cpop = (fx[key + s] for s in ("_scalars", "_pop", "_epray", "_epoints", "_chainpop"))
    out, r = [], 0
    for k in range(len(sc)):
        rays = []
        for j in range(ep.shape[1]):
            m = ep[k, j]
            pts, pp = epts[r:r + m + 1], cpop[r:r + m + 1]
            r += m + 1
            st = np.sqrt((pts[:-1, 0] - pts[1:, 0]) ** 2 + (pts[:-1, 1] - pts[1:, 1]) ** 2 +
                         (pts[:-1, 2] - pts[1:, 2]) ** 2)
            rays.append((pp, st))
        out.append((sc[k], pop[k], rays))
    return out


def rel_err(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(b) & ~((a == 0) & (b == 0))
    if not ok.any():
        return 0.0
    return float(np.max(np.abs(a[ok] - b[ok]) / np.maximum(np.abs(b[ok]), 1e-300)))


@pytest.mark.parametrize("name", [c[0] for c in lamda.COOLANTS])
@pytest.mark.parametrize("variant", ["3d", "1d"])
def test_lvg_matches_reference_fixtures(tables, fx, name, variant):
    key = f"{name}_{variant}"
    tab = tables[name]
    for k, (sc, pop, rays) in enumerate(fixture_cases(fx, key)):
        T, Td, rho, met, vt, g2d, xs = sc[:7]
        res = oracle.lvg_point(tab, T, Td, rho, met, vt, g2d, sc[7:] * rho, xs, pop, rays,
                               pseudo_ray=6 if variant == "1d" else -1)
        assert rel_err(res["C"], fx[key + "_C"][k]) <= 1e-14
        assert rel_err(res["transition"], fx[key + "_transition"][k]) <= 1e-14
        for f in ("cooling", "line", "tau", "beta"):
            assert rel_err(res[f], fx[key + "_" + f][k]) <= 1e-13, (f, k)
        ref = fx[key + "_solution"][k]
        assert np.max(np.abs(res["solution"] - ref)) <= 1e-12 * ref.sum(), k
        # the solver alone, on the reference's own transition matrix: bitwise
        tr = np.ascontiguousarray(fx[key + "_transition"][k].ravel(order="F"))
        sol = np.zeros(tab.cool.nlev)
        oracle._thermlib().pdr_solvlevpop(tab.cool.nlev, tr, float(xs * rho), sol)
        assert np.array_equal(sol, ref), k


def test_dust_temperature_matches_reference(fx):
    got = np.array([oracle.dust_temperature(u, r) for u, r in zip(fx["dust_uv"], fx["dust_rad_surface"])])
    assert np.array_equal(got, fx["dust_T"])
    assert got.min() >= 10.0


@pytest.mark.skipif(not (os.path.isdir(REF) and oracle.have_ref_lvg()), reason="reference build absent")
def test_coolant_reader_matches_readinput():
    for name, fn, nlev, ntemp in lamda.COOLANTS:
        c = lamda.read_coolant(os.path.join(REF, fn), nlev, ntemp, name)
        ref = oracle.ref_coolant_tables(os.path.join(REF, fn), nlev, ntemp)
        for a, b in zip([c.energies, c.weights, c.A, c.B, c.freq, c.temperatures, c.col], ref):
            assert rel_err(a, b) <= 1e-15
        # the stored tables are the reader's
        s = lamda.load_coolant(name)
        for k in ("energies", "weights", "A", "B", "freq", "temperatures", "col"):
            assert np.array_equal(getattr(s, k), getattr(c, k))


def test_lte_closed_form(tables):
    for name, tab in tables.items():
        c = tab.cool
        for T in (5.0, 50.0, 500.0, 5000.0):
            p = oracle.lte_populations(tab, T, 1.0e3)
            assert abs(p.sum() - 1.0e3) <= 1e-12 * 1.0e3
            i = 1
            expect = c.weights[i] / c.weights[0] * np.exp(-(c.energies[i] - c.energies[0]) / (lamda.KB * T))
            assert p[i] / p[0] == pytest.approx(expect, rel=1e-12)


@pytest.mark.parametrize("name", ["CII", "CI", "OI", "CO"])
def test_collision_dominated_limit_is_lte(tables, name):
    """far above the critical density the level populations solvlevpop
    returns are the LTE populations at the gas temperature.  Exactly so at a
    temperature of the rate table (the excitation rates are set by detailed
    balance at each tabulated temperature and interpolated linearly in
    between), with ortho/para-H2 as the only partners."""
    tab = tables[name]
    tt = tab.cool.temperatures
    common = np.intersect1d(tt[1][tt[1] > 0], tt[2][tt[2] > 0])
    T = float(common[common >= 20.0][0])
    rho, x = 1.0e12, 1.0e-4
    dens = np.array([0.0, 0.0, 0.0, 0.0, 0.5]) * rho
    lte = oracle.lte_populations(tab, T, x * rho)
    rays = [(lte[None, :], np.zeros(0)) for _ in range(12)]  # no evaluation points: optically thin
    res = oracle.lvg_point(tab, T, 2.7, rho, 1.0, 1.0e5, 100.0, dens, x, lte, rays)
    big = lte > 1e-6 * lte.sum()
    assert np.allclose(res["solution"][big], lte[big], rtol=1e-5), (res["solution"][big], lte[big])
    # no evaluation points: tau 0 and escape probability 1 on every ray of
    # every pair whose upper level is populated
    n = tab.cool.nlev
    for il in range(n):
        if lte[il] > 0:
            for jl in range(il):
                assert np.all(res["tau"][il, jl, :] == 0.0) and np.all(res["beta"][il, jl, :] == 1.0)


def _search(T0, balance, Tmin=2.7, Tmax=30000.0, Fcrit=0.005, temp_fix=True, max_steps=200):
    """the oracle's temperature search (thermal_loop.ThermalLoop.update_temperature)
    on one point whose heating minus cooling is balance(T) -> (H, C)"""
    from thermal_loop import ThermalLoop
    L = ThermalLoop.__new__(ThermalLoop)
    L.Tmin, L.Tmax, L.Fcrit, L.Tdiff, L.temp_fix = Tmin, Tmax, Fcrit, 0.01, temp_fix
    L.T = np.array([T0])
    L.Tprev = L.T.copy()
    L.Tlow = np.maximum(L.T / 2.0, Tmin)
    L.Thigh = np.minimum(L.T * 1.5, Tmax)
    L.prev = [""]
    L.binarychop = np.zeros(1, bool)
    L.converged = np.zeros(1, bool)
    L.doleveltmin = np.zeros(1, bool)
    L.level_converged = np.zeros(1, bool)
    L.first_time = True
    trail = []
    for k in range(max_steps):
        Tnow = L.T[0]
        h, c = balance(Tnow)
        fmean = h - c
        L.update_temperature(0, fmean, 2.0 * abs(fmean) / abs(h + c))
        L.first_time = False
        trail.append((Tnow, L.T[0], L.Tlow[0], L.Thigh[0], bool(L.binarychop[0])))
        if L.converged[0]:
            return L, trail
    raise AssertionError("no convergence")


@pytest.mark.parametrize("T0,Tstar", [(10.0, 137.0), (500.0, 41.5), (80.0, 80.6), (30.0, 3000.0)])
def test_temperature_search_brackets_the_balance(T0, Tstar):
    """heating a, cooling b T^2: the 30% steps bracket T* = sqrt(a/b), the
    binary chop keeps it bracketed, and the search stops where the relative
    imbalance 2|H-C|/(H+C) is within Fcrit -- the temperature the point
    reports (previousgastemperature) within Fcrit/2 of T*."""
    a = 1.0e-20
    b = a / Tstar ** 2
    L, trail = _search(T0, lambda T: (a, b * T * T))
    for Told, Tnew, lo, hi, chop in trail:
        if chop:
            assert lo <= Tnew <= hi
    # once chopping, the bracket holds T*
    chopped = [t for t in trail if t[4]]
    assert chopped and all(lo <= Tstar * (1 + 1e-12) and hi >= Tstar * (1 - 1e-12) for _, _, lo, hi, _ in chopped)
    Trep = L.Tprev[0]
    assert abs(Trep - Tstar) / Tstar <= 0.005 / 2 * 1.01, (Trep, Tstar)


def test_temperature_search_stops_at_the_bounds():
    """balance below Tmin: the 30% cooling steps stop at Tmin, where the
    temperature no longer moves and TEMP_FIX ends the search; above Tmax the
    search ends at Tmax"""
    a = 1.0e-20
    L, trail = _search(20.0, lambda T: (a, a * (T / 1.0) ** 2), max_steps=400)
    assert L.T[0] == 2.7 and L.Tprev[0] == 2.7 and not L.doleveltmin[0]
    assert [t[1] for t in trail[:3]] == [14.0, 0.7 * 14.0, 0.7 * (0.7 * 14.0)]
    L, trail = _search(20000.0, lambda T: (a, a * (T / 1.0e5) ** 2))
    assert L.T[0] == 30000.0 and L.Tprev[0] == 30000.0
    # without TEMP_FIX a point held at Tmin never converges
    with pytest.raises(AssertionError, match="no convergence"):
        _search(20.0, lambda T: (a, a * (T / 1.0) ** 2), temp_fix=False)
