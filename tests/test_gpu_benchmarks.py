"""End to end: the reference's own benchmark models B01-B04 (Benchmarks.tgz,
its PSEUDO_1D reduced-network runs, kept as tests/golden/benchmarks_fin.npz
with the input point files 1Dn40.dat and 1Dn20.dat from ics.tgz) run by the
model driver (pdrchem.model.run: rays, columns, rates, chemistry, heating,
level populations and the thermal balance, all on the device) and compared
with the reference's output files value by value.

The parameters of each model are read off its output: n_H 1e4 (1Dn40) or
1e2 (1Dn20), the surface field G0 from the UV column, zeta from the cosmic-ray
heating column (9.4 eV * 1.3e-17 zeta n x(H2)), v_turb 1 km/s from the
turbulent heating, Tmin 10 K from the floor of the temperature profiles; the
rest as params.dat.

Bar.  B01, B02, B04: every value of .pdr.fin, .cool.fin and .heat.fin within
one unit of the fourth printed digit (they come out identical in practice:
the test reports the count), the .line.fin lines above 1e-6 of their column's
largest likewise.  B03 (zeta 100x): its thermal search lands a few points on
a neighbouring temperature: gas temperatures within 1%, abundances above
1e-12 within 2%, cooling and heating terms above 1e-3 of the point's total
within 10%.
"""
import os
import re

import numpy as np
import pytest

from pdrchem import lib, model

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = {"B01": ("1Dn40.dat", 1.0, 1.3e-17), "B02": ("1Dn20.dat", 1.0, 1.3e-17), "B03": ("1Dn40.dat", 1.0, 1.3e-15),
         "B04": ("1Dn40.dat", 1000.0, 1.3e-17)}


def load_fin(path):
    """a .fin file; three-digit exponents are printed without the 'E'"""
    with open(path) as f:
        return np.array([[float(re.sub(r"(\d)([+-]\d{3})$", r"\1E\2", t)) for t in line.split()] for line in f])


def last_digit_ok(got, ref):
    """|got - ref| within one unit of ref's fourth significant digit"""
    unit = np.where(ref != 0, 10.0 ** (np.floor(np.log10(np.abs(np.where(ref != 0, ref, 1.0)))) - 3), 0.0)
    return np.abs(got - ref) <= unit * (1 + 1e-9)


@pytest.fixture(scope="module")
def ref():
    return np.load(os.path.join(GOLD, "benchmarks_fin.npz"))


@pytest.mark.parametrize("case", list(CASES))
def test_benchmark_model_reproduces_reference_outputs(case, ref, tmp_path):
    inp, G0, zeta = CASES[case]
    p = model.Params(input=inp, output=case, Gext=G0, zeta=zeta, Tmin=10.0)
    print(f"benchmark {case}: {inp}, G0={G0}, zeta={zeta}", flush=True)
    r = model.run(p, directory=GOLD, network="reduced", pseudo_1d=True, precision=lib.MIXED,
                  out_dir=str(tmp_path), log=None)
    r["cloud"].close()
    assert r["converged"]
    print(f"  converged after {r['iterations']} iterations, {r['seconds']}", flush=True)
    got = {k: load_fin(str(tmp_path / f"{case}.{k}.fin")) for k in ("pdr", "cool", "heat", "line")}
    want = {k: ref[f"{case}_{k}"] for k in ("pdr", "cool", "heat", "line")}
    for k in got:
        assert got[k].shape == want[k].shape, k
        assert np.array_equal(got[k][:, 0], want[k][:, 0]), k          # ids and row order
    # positions, Av, dust temperature, density and field do not depend on the search
    pd, pw = got["pdr"], want["pdr"]
    for col in (1, 2, 4, 5, 6, 7):
        assert np.all(last_digit_ok(pd[:, col], pw[:, col])), col
    line_sig = np.abs(want["line"]) >= 1e-6 * np.abs(want["line"]).max(axis=0, keepdims=True)
    if case != "B03":
        for k in ("pdr", "cool", "heat"):
            ok = last_digit_ok(got[k], want[k])
            print(f"  .{k}.fin: {int(np.sum(got[k] == want[k]))} of {got[k].size} values identical", flush=True)
            assert np.all(ok), (k, np.argwhere(~ok)[:5])
        ok = last_digit_ok(got["line"], want["line"]) | ~line_sig
        assert np.all(ok), np.argwhere(~ok)[:5]
    else:
        Tg, Tw = pd[:, 3], pw[:, 3]
        assert np.max(np.abs(Tg - Tw) / Tw) <= 0.01
        ab, aw = pd[:, 8:], pw[:, 8:]
        m = aw > 1e-12
        assert np.max(np.abs(ab[m] - aw[m]) / aw[m]) <= 0.02
        for k, tot in (("cool", 7), ("heat", 14)):
            g, w = got[k][:, 3:tot + 1], want[k][:, 3:tot + 1]
            big = np.abs(w) >= 1e-3 * np.abs(want[k][:, tot:tot + 1])
            assert np.max(np.abs(g[big] - w[big]) / np.abs(w[big])) <= 0.10, k
        print(f"  B03: {int(np.sum(~last_digit_ok(pd, pw)))} .pdr.fin values beyond the last printed digit, "
              f"largest temperature difference {np.max(np.abs(Tg - Tw) / Tw):.2%}", flush=True)

