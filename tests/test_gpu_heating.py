"""Device heating terms (pdrchem_heating_device, CALC_HEATING of
sub_calculate_heating.F90:6-386) against the reference's own values (golden
vectors from the amdflang build, tests/golden/make_heating_fixtures.py) and
against the oracle (oracle/heating.c) on a larger seeded batch."""
import os

import numpy as np
import pytest

import oracle as O
from pdrchem import lib
from pdrchem.netgen import load_network

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
NET_ID = {"reduced": 0, "full": 1, "mynetwork": 2}
# device pow/exp/sqrt against the host libm, and the chemical channels summed
# in another order (all positive): a few ulps per term
RTOL = 1e-12


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _run(name, n, T, Td, uv, ab, rate):
    idx = O.heating_indices(load_network(name)) + 1
    print(f"heating_device {name} npts={len(n)}", flush=True)
    h = lib.heating_device(NET_ID[name], idx, *[_dev(v) for v in (n, T, Td, uv, ab, rate)])
    torch.cuda.synchronize()
    return h.cpu().numpy()


@pytest.mark.parametrize("name", ["reduced", "full"])
def test_heating_device_matches_reference_fixtures(name):
    z = np.load(os.path.join(HERE, "golden", f"heating_{name}.npz"))
    h = _run(name, z["n"], z["T"], z["Td"], z["uv"], z["ab"], z["rate"])
    assert np.allclose(h, z["heat"], rtol=RTOL, atol=0)


@pytest.mark.parametrize("name,npts", [("reduced", 20000), ("full", 3000), ("mynetwork", 5000)])
def test_heating_device_matches_oracle(name, npts):
    rng = np.random.default_rng(5)
    net = load_network(name)
    o = O.Oracle(name)
    T = 10 ** rng.uniform(1, 4, npts)
    Td = rng.uniform(5, 150, npts)
    # rates at a handful of conditions, tiled: the heating reads them as given
    base = o.rates(T[:16], Td[:16], np.ones((16, 12)), rng.uniform(0, 10, (16, 12)))
    rate = base[rng.integers(0, 16, npts)]
    ab = 10 ** rng.uniform(-12, -1, (npts, net.nspec))
    n = 10 ** rng.uniform(1, 6, npts)
    uv = 10 ** rng.uniform(-3, 5, npts)
    h = _run(name, n, T, Td, uv, ab, rate)
    sel = rng.choice(npts, 400, replace=False)
    ref = O.heating(name, net, n[sel], T[sel], Td[sel], uv[sel], ab[sel], rate[sel])
    assert np.allclose(h[sel], ref, rtol=RTOL, atol=0)
    assert np.all(np.isfinite(h))


def test_heating_device_rejects_bad_input():
    z = np.load(os.path.join(HERE, "golden", "heating_reduced.npz"))
    idx = O.heating_indices(load_network("reduced")) + 1
    args = [_dev(z[k]) for k in ("n", "T", "Td", "uv", "ab", "rate")]
    bad = idx.copy()
    bad[0] = 0                                   # slots are 1-based
    with pytest.raises(lib.PdrChemError):
        lib.heating_device(0, bad, *args)
    with pytest.raises(lib.PdrChemError):
        lib.heating_device(3, idx, *args)        # no such network
