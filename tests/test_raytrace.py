"""Ray tracer of the oracle (oracle/raytrace.c) against the reference's own
evaluation_points (golden vectors written by tests/golden/make_raytrace_fixtures.py
from the amdflang build of eval_points.F90, healpix.F90, heapsort.F90), and
against what geometry fixes: HEALPix pixel centres lie in their own pixel,
the rays are unit vectors, a uniform-density column is the density times the
path to the last evaluation point, and the UNI surface field falls on the one
ray most opposed to the field."""
import os

import numpy as np
import pytest

import oracle as O
from pdrchem import lib

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "raytrace_ref.npz"))
MAXP = 40


@pytest.mark.parametrize("name", ["jitter", "grid", "grid_ion"])
def test_heapsort_order_matches_reference_bitwise(name):
    """With the reference's heapsort order the restatement finds the same
    evaluation points, in the same order, at the same coordinates (bitwise)."""
    xyz, et = GOLD[f"{name}_xyz"], GOLD[f"{name}_etype"]
    epray, proj, epoint = O.eval_points(xyz, order=O.ORDER_HEAPSORT, maxpoints=MAXP, etype=et)
    assert np.array_equal(epray, GOLD[f"{name}_epray"])
    assert np.array_equal(proj, GOLD[f"{name}_proj"])
    assert np.array_equal(epoint, GOLD[f"{name}_epoint"])


def test_healpix_vectors_match_reference():
    assert np.array_equal(O.hp_vectors(0), GOLD["vectors"])


@pytest.mark.parametrize("level", [0, 1, 2])
def test_healpix_centres_in_own_pixel(level):
    v = O.hp_vectors(level)
    assert v.shape == (12 << (2 * level), 3)
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-10)
    for j in range(v.shape[0]):
        assert O.hp_vec2pix(level, v[j]) == j


def test_healpix_equal_areas():
    """HEALPix pixels have equal areas: isotropic directions fall evenly."""
    rng = np.random.default_rng(3)
    d = rng.normal(size=(24000, 3))
    counts = np.bincount([O.hp_vec2pix(0, r) for r in d], minlength=12)
    assert counts.min() > 1700 and counts.max() < 2300


def test_stable_order_equals_heapsort_without_ties():
    """On a cloud without distance ties the (distance, index) order of the
    device tracer gives the reference's evaluation points, except on rays
    touched by heapsort's last-element quirk (the highest-id point is left out
    of the initial heap, so when it is the farthest it is visited second to
    last)."""
    xyz = GOLD["jitter_xyz"]
    e_h, p_h, _ = O.eval_points(xyz, order=O.ORDER_HEAPSORT, maxpoints=MAXP)
    e_s, p_s, _ = O.eval_points(xyz, order=O.ORDER_STABLE, maxpoints=MAXP)
    same = (e_h == e_s) & np.all(p_h == p_s, axis=2)
    n = len(xyz)
    for p, j in np.argwhere(~same):
        d = np.linalg.norm(xyz - xyz[p], axis=1)
        d[p] = -1
        far = np.argsort(d)[-2:]
        last = n - 1 if p != n - 1 else n - 2
        assert last in far, (p, j)
    assert same.mean() > 0.99


def test_uniform_density_column_is_path_length():
    """rho uniform: the trapezoid sum telescopes to rho * |last epoint - origin|
    (evaluation points advance monotonically along the ray)."""
    xyz = GOLD["grid_xyz"]
    e, p, ep = O.eval_points(xyz, order=O.ORDER_STABLE, maxpoints=MAXP)
    rho = np.full(len(xyz), 250.0)
    xs = np.full((len(xyz), 1), 1e-4)
    Av, cols = O.ray_columns(e, p, ep, rho, xsel=xs)
    n, nr = e.shape
    last = ep[np.arange(n)[:, None], np.arange(nr)[None, :], e]
    L = np.linalg.norm(last - xyz[:, None, :], axis=2)
    assert np.allclose(Av, 250.0 * O.PC_CM * L * O.AV_FAC, rtol=1e-12, atol=0)
    assert np.allclose(cols[:, :, 0], 250.0 * 1e-4 * O.PC_CM * L, rtol=1e-12, atol=0)
    assert np.all(Av[e == 0] == 0)


def test_uniform_density_column_same_in_both_orders():
    """Tie order changes which points are sampled, not the path: with uniform
    density both orders give the same Av on a grid full of ties."""
    xyz = GOLD["grid_xyz"]
    rho = np.full(len(xyz), 1e3)
    A = [O.ray_columns(*O.eval_points(xyz, order=o, maxpoints=MAXP), rho)[0]
         for o in (O.ORDER_HEAPSORT, O.ORDER_STABLE)]
    assert np.allclose(A[0], A[1], rtol=1e-13, atol=0)


def test_ionised_point_stops_ray():
    xyz, et = GOLD["grid_ion_xyz"], GOLD["grid_ion_etype"]
    e, p, _ = O.eval_points(xyz, order=O.ORDER_STABLE, maxpoints=MAXP, etype=et)
    ion = int(np.argmax(et == 2))
    n, nr = e.shape
    hits = 0
    for q in range(n):
        for j in range(nr):
            ids = list(p[q, j, 1:e[q, j] + 1])
            if ion in ids:
                hits += 1
                assert ids[-1] == ion
    assert hits > 0


def test_uvfield_uni_and_guess_temperature():
    v = O.hp_vectors(0)
    Av = np.zeros((3, 12))
    Av[1] = 1.0
    G, uv, T = O.uvfield(Av, gext=(100.0, 0.0, 0.0))
    jmin = int(np.argmin(v[:, 0]))                    # ray pointing most towards -x
    assert np.all(G[:, [j for j in range(12) if j != jmin]] == 0)
    assert np.allclose(G[:, jmin], -100.0 * v[jmin, 0], rtol=1e-15)
    assert np.isclose(uv[0], G[0, jmin]) and np.isclose(uv[1], G[1, jmin] * np.exp(-O.UV_FAC))
    assert np.allclose(T, 10.0 * (1.0 + (2.0 * uv) ** (1.0 / 3.0)), rtol=1e-15)
    Giso, uviso, _ = O.uvfield(Av[:1], gext=(12.0, 0.0, 0.0), fieldchoice=0)
    assert np.allclose(Giso, 1.0) and np.isclose(uviso[0], 12.0)


@pytest.mark.skipif(not O.have_ref_raytrace(), reason="reference ray tracer not built (needs /root/reference)")
def test_live_reference_on_random_cloud():
    """A fresh random cloud through the reference build (not a fixture)."""
    rng = np.random.default_rng(11)
    xyz = rng.uniform(0, 1, (150, 3))
    vec, e_r, p_r, ep_r = O.ref_eval_points(xyz, maxpoints=MAXP)
    e_o, p_o, ep_o = O.eval_points(xyz, order=O.ORDER_HEAPSORT, maxpoints=MAXP)
    assert np.array_equal(e_r, e_o) and np.array_equal(p_r, p_o) and np.array_equal(ep_r, ep_o)


def _library_chains(xyz, order, etype=None, maxp=MAXP):
    """The library's host tracer in the oracle's padded layout."""
    count, offs, proj, step = lib.trace_points(xyz, order=order, etype=etype)
    n, nr = count.shape
    P = np.zeros((n, nr, maxp + 1), dtype=np.int32)
    S = np.zeros((n, nr, maxp + 1))
    for q in range(n):
        for j in range(nr):
            o, k = offs[q * nr + j], count[q, j]
            P[q, j, 0] = q
            P[q, j, 1:k + 1] = proj[o:o + k]
            S[q, j, 1:k + 1] = step[o:o + k]
    return count, P, S


@pytest.mark.parametrize("name", ["jitter", "grid", "grid_ion"])
def test_library_tracer_reference_order_matches_reference_bitwise(name):
    """The product's host tracer (libpdrchem.so, PDRCHEM_ORDER_REFERENCE) finds
    the reference's evaluation points, and its step lengths are those of the
    reference's evaluation-point coordinates, bitwise."""
    xyz, et = GOLD[f"{name}_xyz"], GOLD[f"{name}_etype"]
    count, P, S = _library_chains(xyz, lib.ORDER_REFERENCE, etype=et)
    assert np.array_equal(count, GOLD[f"{name}_epray"])
    assert np.array_equal(P, GOLD[f"{name}_proj"])
    ep = GOLD[f"{name}_epoint"]
    d = ep[:, :, 1:, :] - ep[:, :, :-1, :]
    ref_step = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
    mask = np.arange(1, MAXP + 1)[None, None, :] <= count[:, :, None]
    assert np.array_equal(S[:, :, 1:][mask], ref_step[mask])


@pytest.mark.parametrize("name", ["jitter", "grid"])
def test_library_tracer_distance_order_matches_oracle(name):
    xyz = GOLD[f"{name}_xyz"]
    count, P, S = _library_chains(xyz, lib.ORDER_DISTANCE)
    e, p, ep = O.eval_points(xyz, order=O.ORDER_STABLE, maxpoints=MAXP)
    assert np.array_equal(count, e) and np.array_equal(P, p)
