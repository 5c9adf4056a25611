"""Device thermal balance against the reference and the oracle.

* pdrchem_level_populations_device on the reference's own escape_probability
  + solvlevpop results (tests/golden/lvg_fixtures.npz, written by the
  reference's Fortran built unchanged): every coolant, the 3D and the
  PSEUDO_1D build, including its special paths (a ray without evaluation
  points, an empty level, a pair at its weights' ratio, masing chains, a NaN
  cooling the reference itself produces);
* pdrchem_dust_temperature_device on the reference's dust temperatures;
  pdrchem_lte_populations_device on the oracle's LTE populations;
* the whole loop (pdrchem_cloud_thermal_iterate) in lockstep with the
  oracle's restatement of 3DPDR.F90:590-1241 (oracle/thermal_loop.py) on a
  small cloud until every point is thermally converged: after every
  iteration the loop's counters and flags and the temperatures agree
  exactly, level populations, line cooling and heating to rounding.
"""
import os

import numpy as np
import pytest

import oracle as O
from pdrchem import cloud as CL
from pdrchem import lamda, lib, workloads
from pdrchem.netgen import load_network
from thermal_loop import ThermalLoop

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# the device forms each optical depth from level columns summed once per ray
# (the reference sums per pair along the chain): equal to rounding, which the
# solve of the level populations carries through its condition number
SOL_TOL = 1e-11    # of the coolant's total density
COOL_RTOL = 1e-11
TAU_RTOL = 1e-11


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).cuda()


def nan_close(a, b, rtol, atol=0.0):
    a, b = np.asarray(a), np.asarray(b)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    m = ~np.isnan(b)
    return np.all(np.abs(a[m] - b[m]) <= rtol * np.abs(b[m]) + atol)


@pytest.fixture(scope="module")
def coolants():
    return lamda.load_coolants()


@pytest.fixture(scope="module")
def tables(coolants):
    return lib.CoolantTables(coolants, np.arange(9, dtype=np.int32))


def fixture_batch(fx, key, ci, nlevtot, off):
    """the fixture cases of one coolant as one device batch: point k's row
    holds its populations, every chain element has a row of its own"""
    sc, pop, ep, epts, cpop = (fx[key + s] for s in ("_scalars", "_pop", "_epray", "_epoints", "_chainpop"))
    nc, nr = ep.shape
    nlev = pop.shape[1]
    rows = [np.zeros(nlevtot) for _ in range(nc)]
    for k in range(nc):
        rows[k][off:off + nlev] = pop[k]
    count = ep.reshape(-1).astype(np.int32)
    offs = np.concatenate([[0], np.cumsum(count)]).astype(np.int64)
    proj, step = [], []
    r = 0
    for k in range(nc):
        for j in range(nr):
            m = ep[k, j]
            pts, pp = epts[r:r + m + 1], cpop[r:r + m + 1]
            r += m + 1
            step.extend(np.sqrt((pts[:-1, 0] - pts[1:, 0]) ** 2 + (pts[:-1, 1] - pts[1:, 1]) ** 2 +
                                (pts[:-1, 2] - pts[1:, 2]) ** 2))
            for i in range(1, m + 1):
                proj.append(len(rows))
                row = np.zeros(nlevtot)
                row[off:off + nlev] = pp[i]
                rows.append(row)
    ab = np.zeros((nc, 9))
    ab[:, ci] = sc[:, 6]
    ab[:, 4:9] = sc[:, 7:12]
    return sc, np.array(rows), count, offs, np.array(proj, dtype=np.int32), np.array(step + [0.0]), ab


@pytest.mark.parametrize("variant", ["3d", "1d"])
@pytest.mark.parametrize("ci,name", list(enumerate(["CII", "CI", "OI", "CO"])))
def test_level_populations_match_reference_fixtures(fx_lvg, tables, coolants, ci, name, variant):
    key = f"{name}_{variant}"
    off = int(tables.offsets[ci])
    sc, rows, count, offs, proj, step, ab = fixture_batch(fx_lvg, key, ci, tables.nlevtot, off)
    nc = len(sc)
    phys = lib.default_phys(v_turb=float(sc[0, 4]), metallicity=float(sc[0, 3]), g2d=float(sc[0, 5]))
    assert np.all(sc[:, 3:6] == sc[0, 3:6])
    print(f"level populations {key}: {nc} points, {len(proj)} evaluation points", flush=True)
    tau_ray = 6 if variant == "1d" else 5
    sol, cool, lines, tau = lib.level_populations_device(
        tables, ci, dev(sc[:, 0]), dev(sc[:, 1]), dev(sc[:, 2]), dev(ab), dev(rows), dev(count, torch.int32),
        dev(offs, torch.int64), dev(proj, torch.int32), dev(step), nrays=12, pseudo_ray=6 if variant == "1d" else -1,
        phys=phys, tau_ray=tau_ray)
    torch.cuda.synchronize()
    sol, cool, lines, tau = sol.cpu().numpy(), cool.cpu().numpy(), lines.cpu().numpy(), tau.cpu().numpy()
    nlev = coolants[ci].nlev
    ref_sol = fx_lvg[key + "_solution"]
    ref_cool = fx_lvg[key + "_cooling"]
    ref_line = fx_lvg[key + "_line"]
    ref_tau = fx_lvg[key + "_tau"]
    ref_pop = fx_lvg[key + "_pop"]
    for k in range(nc):
        tot = ref_sol[k].sum()
        assert np.max(np.abs(sol[k, off:off + nlev] - ref_sol[k])) <= SOL_TOL * tot, (k, sol[k], ref_sol[k])
    assert nan_close(cool[:, ci], ref_cool, COOL_RTOL), (cool[:, ci], ref_cool)
    # the lines of this coolant among the 17
    for il in range(nlev):
        for jl in range(il):
            s = lib_line_slot(ci, il, jl)
            if s >= 0:
                assert nan_close(lines[:, s], ref_line[:, il, jl], COOL_RTOL), (il, jl)
                # optical depth along tau_ray where the reference forms one
                # (a populated upper level off the weights' ratio)
                for k in range(nc):
                    tpop = (ref_pop[k, jl] * coolants[ci].weights[il]) / (ref_pop[k, il] * coolants[ci].weights[jl]) \
                        - 1.0 if ref_pop[k, il] != 0 else 0.0
                    r = ref_tau[k, il, jl, tau_ray] if ref_pop[k, il] != 0 and abs(tpop) >= 1e-50 else 0.0
                    scale = np.max(np.abs(ref_tau[k, il, jl]))
                    assert abs(tau[k, s] - r) <= TAU_RTOL * abs(r) + 1e-14 * scale, (k, il, jl, tau[k, s], r)
    # the other coolants' columns are left alone
    others = [c for c in range(4) if c != ci]
    assert np.all(cool[:, others] == 0.0)


def lib_line_slot(ci, il, jl):
    from thermal_loop import line_slot
    return line_slot(ci, il, jl)


@pytest.fixture(scope="module")
def fx_lvg():
    return np.load(os.path.join(GOLD, "lvg_fixtures.npz"))


def test_dust_temperature_matches_reference(fx_lvg):
    uv, rs = fx_lvg["dust_uv"], fx_lvg["dust_rad_surface"]
    print(f"dust temperature: {len(uv)} points", flush=True)
    td = lib.dust_temperature_device(dev(rs), dev(uv))
    torch.cuda.synchronize()
    assert np.allclose(td.cpu().numpy(), fx_lvg["dust_T"], rtol=1e-13, atol=0)


def test_lte_matches_oracle(tables, coolants):
    rng = np.random.default_rng(11)
    n = 500
    T = 10 ** rng.uniform(0.5, 4, n)
    rho = 10 ** rng.uniform(1, 6, n)
    ab = 10 ** rng.uniform(-9, -3, (n, 9))
    skip = (rng.uniform(size=n) < 0.2).astype(np.uint8)
    print(f"LTE populations: {n} points", flush=True)
    pop0 = np.full((n, tables.nlevtot), -1.0)
    pop = lib.lte_populations_device(tables, dev(T), dev(ab), dev(rho), skip=dev(skip, torch.uint8), pop=dev(pop0))
    torch.cuda.synchronize()
    pop = pop.cpu().numpy()
    assert np.all(pop[skip == 1] == -1.0)
    tabs = [O.CoolantTables(c) for c in coolants]
    for p in np.where(skip == 0)[0]:
        for ci, tab in enumerate(tabs):
            a, b = tables.offsets[ci], tables.offsets[ci + 1]
            ref = O.lte_populations(tab, T[p], ab[p, ci] * rho[p])
            assert np.allclose(pop[p, a:b], ref, rtol=1e-13, atol=1e-300), (p, ci)


def test_level_populations_reject_bad_arguments(tables):
    z = dev(np.zeros(4))
    with pytest.raises(lib.PdrChemError):   # coolant 4 does not exist
        lib.level_populations_device(tables, 4, z, z, z, dev(np.zeros((4, 9))), dev(np.zeros((4, 56))),
                                     dev(np.zeros(48), torch.int32), dev(np.zeros(49), torch.int64), None, None,
                                     nrays=12)
    with pytest.raises(lib.PdrChemError):   # abundance slots beyond nspec
        lib.level_populations_device(tables, 0, z, z, z, dev(np.zeros((4, 5))), dev(np.zeros((4, 56))),
                                     dev(np.zeros(48), torch.int32), dev(np.zeros(49), torch.int64), None, None,
                                     nrays=12)


STAT_KEYS = ("lvg", "thermal_points", "thermally_converged", "done", "iteration", "levpop_iteration", "level_conv",
             "first_time")


@pytest.mark.parametrize("name,pseudo_ray", [("reduced", -1), ("reduced", 6), ("mynetwork", -1)])
def test_thermal_loop_matches_oracle(coolants, name, pseudo_ray):
    shape, H = (4, 3, 3), 0.0625
    net = load_network(name)
    n = int(np.prod(shape))
    rho = 1e3 * np.exp(0.3 * np.random.default_rng(3).normal(size=n))
    c = CL.Cloud(net, shape, H, rho, gext=(10.0, 0.0, 0.0))
    try:
        c.thermal_init(coolants, lib.default_thermal_params(pseudo_ray=pseudo_ray))
        c.set_abundances(workloads.initial_abundances(net, n))
        print(f"thermal loop {name} {shape} pseudo_ray={pseudo_ray}: 3 chemistry iterations", flush=True)
        for _ in range(3):
            st = c.iterate(t_end=1e7, conv_tol=0.0, precision=lib.FP64)
            assert st["failed"] == 0
        count, offs, proj, step = c.evaluation_points()
        L = ThermalLoop(name, net, coolants, rho, c.get(CL.UVFIELD), c.get(CL.G), c.get(CL.AV),
                        c.get(CL.TEMPERATURE), count, offs, proj, step, pseudo_ray=pseudo_ray)
        assert np.allclose(c.get(CL.DUST_TEMPERATURE), L.Td, rtol=1e-13, atol=0)
        chem_steps = updates = 0
        worst = dict(pop=0.0, cool=0.0, heat=0.0)
        for it in range(1, 301):
            before = c.get(CL.ABUNDANCE)
            chem = L.chemistry_due()
            got = c.thermal_iterate(t_end=1e7, precision=lib.FP64)
            ab, cols = c.get(CL.ABUNDANCE), c.get(CL.COLUMNS)
            if chem:
                chem_steps += 1
                assert got["chemistry_solves"] > 0 and got["failed"] == 0
            else:
                assert np.array_equal(ab, before)
            T_before = L.T.copy()
            ref = L.iterate(ab, cols)
            updates += int(np.any(L.T != T_before))
            for k in STAT_KEYS:
                assert got[k] == ref.get(k, 0), (it, k, got, ref)
            if not ref["done"]:
                assert got["level_converged"] == ref["level_converged"], (it, got, ref)
                assert got["coolant_converged"] == ref["coolant_converged"], (it, got, ref)
            assert np.array_equal(c.get(CL.TEMPERATURE), L.T), it
            assert np.array_equal(c.get(CL.PREV_TEMPERATURE), L.Tprev), it
            assert np.array_equal(c.get(CL.THERMAL_CONVERGED) != 0, L.converged), it
            pop = c.get(CL.POPULATIONS)
            for ci in range(4):
                a, b = L.off[ci], L.off[ci + 1]
                tot = np.maximum(L.pop[:, a:b].sum(axis=1, keepdims=True), 1e-300)
                worst["pop"] = max(worst["pop"], float(np.max(np.abs(pop[:, a:b] - L.pop[:, a:b]) / tot)))
            cool = c.get(CL.COOLING)
            scale = np.maximum(np.abs(L.cool).sum(axis=1), 1e-300)
            worst["cool"] = max(worst["cool"], float(np.max(np.abs(cool[:, :4] - L.cool).max(axis=1) / scale)),
                                float(np.max(np.abs(cool[:, 4] - L.total) / scale)))
            heat = c.get(CL.HEATING)
            hs = np.maximum(np.abs(L.heat).max(axis=1, keepdims=True), 1e-300)
            worst["heat"] = max(worst["heat"], float(np.max(np.abs(heat - L.heat) / hs)))
            assert np.allclose(c.get(CL.LINES), L.lines, rtol=1e-8, atol=1e-8 * np.abs(L.lines).max()), it
            assert np.allclose(c.get(CL.OPTICAL_DEPTHS), L.tau_lines, rtol=1e-8,
                               atol=1e-8 * max(np.abs(L.tau_lines).max(), 1e-300)), it
            if ref["done"]:
                break
        print(f"  {it} iterations, {chem_steps} chemistry steps, {updates} temperature updates; worst {worst}",
              flush=True)
        assert ref["done"] and chem_steps >= 2 and updates >= 2
        assert worst["pop"] <= 1e-9 and worst["cool"] <= 1e-9 and worst["heat"] <= 1e-10, worst
        # once done, further calls change nothing
        T = c.get(CL.TEMPERATURE)
        assert c.thermal_iterate(t_end=1e7)["done"] == 1
        assert np.array_equal(c.get(CL.TEMPERATURE), T)
    finally:
        c.close()
