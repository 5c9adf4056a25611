"""Heating terms of the oracle (oracle/heating.c, a restatement of CALC_HEATING,
sub_calculate_heating.F90:6-386) against the reference's own CALC_HEATING:
golden vectors written by tests/golden/make_heating_fixtures.py from the
amdflang build of the reference, and a live run of that build when present."""
import os

import numpy as np
import pytest

import oracle as O
from pdrchem.netgen import load_network

HERE = os.path.dirname(os.path.abspath(__file__))
# the restatement evaluates the same expressions in the same order as the
# reference; differences are libm pow/exp ulps only
RTOL = 1e-13


def _load(name):
    return np.load(os.path.join(HERE, "golden", f"heating_{name}.npz"))


@pytest.mark.parametrize("name", ["reduced", "full"])
def test_heating_matches_reference_fixtures(name):
    z = _load(name)
    h = O.heating(name, load_network(name), z["n"], z["T"], z["Td"], z["uv"], z["ab"], z["rate"])
    assert np.all(np.isfinite(z["heat"]))
    assert np.allclose(h, z["heat"], rtol=RTOL, atol=0)


@pytest.mark.parametrize("name", ["reduced", "full"])
def test_heating_total_and_signs(name):
    """The total is the sum of every term except the photoelectric and the
    Weingartner & Draine ones; cosmic-ray and turbulent heating are positive;
    gas-grain exchange has the sign of T_dust - T."""
    z = _load(name)
    h = z["heat"]
    assert np.allclose(h[:, 11], h[:, [1, 3, 4, 5, 6, 7, 8, 9, 10]].sum(1), rtol=1e-12, atol=0)
    assert np.all(h[:, 7] > 0) and np.all(h[:, 8] > 0)
    assert np.array_equal(np.sign(h[:, 10]), np.sign(z["Td"] - z["T"]))


def test_heating_linear_terms_scale_with_density():
    """C ionization, H2 photodissociation, cosmic-ray and turbulent heating are
    linear in n_H at fixed abundances and rates; H2 formation is quadratic."""
    z = _load("reduced")
    net = load_network("reduced")
    sl = slice(0, 8)
    a = O.heating("reduced", net, z["n"][sl], z["T"][sl], z["Td"][sl], z["uv"][sl], z["ab"][sl], z["rate"][sl])
    b = O.heating("reduced", net, 3 * z["n"][sl], z["T"][sl], z["Td"][sl], z["uv"][sl], z["ab"][sl], z["rate"][sl])
    assert np.allclose(b[:, [3, 5, 7, 8]], 3 * a[:, [3, 5, 7, 8]], rtol=1e-14)
    assert np.allclose(b[:, 4], 9 * a[:, 4], rtol=1e-14)


@pytest.mark.parametrize("name", ["reduced", "full"])
@pytest.mark.skipif(not (O.have_ref_heating("reduced") and O.have_ref_heating("full")),
                    reason="reference heating not built (needs /root/reference)")
def test_heating_live_reference(name):
    rng = np.random.default_rng(77)
    o = O.Oracle(name)
    npts = 6
    T = 10 ** rng.uniform(1, 4, npts)
    Td = rng.uniform(5, 80, npts)
    rate = o.rates(T, Td, np.ones((npts, 12)), rng.uniform(0, 5, (npts, 12)))
    ab = 10 ** rng.uniform(-12, -1, (npts, o.nspec))
    n = 10 ** rng.uniform(1, 6, npts)
    uv = 10 ** rng.uniform(-3, 4, npts)
    ref = O.ref_heating(name, n, T, Td, uv, ab, rate)
    assert np.allclose(O.heating(name, load_network(name), n, T, Td, uv, ab, rate), ref, rtol=RTOL, atol=0)


def _chemical_channels(name):
    """{reaction number: heating per unit rate} of the oracle's chemical term,
    found by switching on one reaction at a time (every other rate zero)"""
    net = load_network(name)
    idx = O.heating_indices(net)
    ab = np.full((1, net.nspec), 1.0e-3)
    n, T, Td, uv = np.array([100.0]), np.array([50.0]), np.array([20.0]), np.array([1.0])
    out = {}
    for k in range(net.nreac):
        rate = np.zeros((1, net.nreac))
        rate[0, k] = 1.0
        chem = O.heating(name, net, n, T, Td, uv, ab, rate)[0, 9]
        if chem != 0.0:
            r = net.reactions[k]
            out[k + 1] = (tuple(sorted(x for x in r.reactants if x)), tuple(sorted(x for x in r.products if x)), chem)
    assert idx[9] >= 0
    return out


def test_mynetwork_chemical_channels_match_the_full_networks():
    """MYNETWORK's pre-set chemical-heating channels (sub_calculate_heating.F90:
    254-259, behind the reference's STOP) are the full network's reactions
    under MYNETWORK's own numbers: each channel's reaction has the same
    reactants and products as a full-network channel and the same energy, and
    the two networks have the same number of channels.  The full network's
    channels are pinned against the reference build (the fixtures above), so
    a wrong reaction number or energy in MYNETWORK's list fails here."""
    full = _chemical_channels("full")
    mine = _chemical_channels("mynetwork")
    assert sorted(mine) == [67, 172, 173, 176, 178, 179, 185, 297, 364]
    by_reaction = {(a, b): e for a, b, e in full.values()}
    assert len(by_reaction) == len(full) == len(mine)
    for k, (a, b, e) in mine.items():
        assert (a, b) in by_reaction, (k, a, b)
        assert e == by_reaction[(a, b)], k
