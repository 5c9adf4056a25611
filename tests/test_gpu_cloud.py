"""Device ray tracer, columns and chemistry iteration of a grid cloud
(pdrchem_cloud_*) against the oracle: evaluation points bitwise equal to the
oracle's (distance, index) order, Av / surface field / temperature / columns
to rounding, and one chemistry iteration within the 1e-4 parity bar of the
FP64 oracle fed with the oracle's own rates."""
import numpy as np
import pytest

import oracle as O
from pdrchem import cloud as CL
from pdrchem import lib, workloads
from pdrchem.netgen import load_network

pytestmark = pytest.mark.gpu
H = 0.0625


def grid_xyz(nx, ny, nz):
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.stack([ix, iy, iz], -1).reshape(-1, 3).astype(float) * H


def smooth_density(nx, ny, nz, seed=5):
    rng = np.random.default_rng(seed)
    xyz = grid_xyz(nx, ny, nz) / H
    c = np.array([nx, ny, nz]) / 2.0
    return 1e3 * np.exp(0.3 * np.sin(xyz[:, 0] / 2.0) + 0.2 * np.cos(xyz[:, 1] / 3.0)
                        - 0.05 * np.sum((xyz - c) ** 2, axis=1) / max(nx, ny, nz) + 0.05 * rng.normal(size=len(xyz)))


def oracle_chains(shape):
    xyz = grid_xyz(*shape)
    e, p, ep = O.eval_points(xyz, order=O.ORDER_STABLE, maxpoints=400)
    return xyz, e, p, ep


@pytest.mark.parametrize("shape", [(6, 7, 8), (9, 5, 4)])
def test_evaluation_points_match_oracle(shape):
    print(f"cloud evaluation points {shape}", flush=True)
    net = load_network("reduced")
    n = int(np.prod(shape))
    c = CL.Cloud(net, shape, H, np.full(n, 1e3))
    count, offs, proj, step = c.evaluation_points()
    xyz, e, p, ep = oracle_chains(shape)
    assert np.array_equal(count, e)
    nr = e.shape[1]
    for q in range(n):
        for j in range(nr):
            o = offs[q * nr + j]
            k = e[q, j]
            assert np.array_equal(proj[o:o + k], p[q, j, 1:k + 1]), (q, j)
            d = np.linalg.norm(np.diff(ep[q, j, :k + 1], axis=0), axis=1)
            assert np.allclose(step[o:o + k], d, rtol=1e-15, atol=0), (q, j)
    c.close()


def test_fields_and_columns_match_oracle():
    shape = (7, 6, 5)
    print(f"cloud fields {shape}", flush=True)
    net = load_network("reduced")
    n = int(np.prod(shape))
    rho = smooth_density(*shape)
    c = CL.Cloud(net, shape, H, rho, gext=(100.0, 0.0, 0.0))
    x0 = workloads.initial_abundances(net, n)
    x0 *= np.exp(0.1 * np.random.default_rng(2).normal(size=x0.shape))
    c.set_abundances(x0)
    xyz, e, p, ep = oracle_chains(shape)
    slots = CL.shield_slots(net)
    xs = np.where(slots[None, :] >= 0, x0[:, np.maximum(slots, 0)], 0.0)
    Av, cols = O.ray_columns(e, p, ep, rho, xsel=xs)
    G, uv, T = O.uvfield(Av, gext=(100.0, 0.0, 0.0))
    assert np.allclose(c.get(CL.AV), Av, rtol=1e-14, atol=0)
    assert np.allclose(c.get(CL.COLUMNS), cols, rtol=1e-14, atol=0)
    assert np.array_equal(c.get(CL.G), G)
    assert np.allclose(c.get(CL.UVFIELD), uv, rtol=1e-13, atol=1e-300)
    assert np.allclose(c.get(CL.TEMPERATURE), T, rtol=1e-13, atol=0)
    c.close()


def test_iteration_matches_oracle_and_skips_converged():
    shape = (5, 5, 4)
    print(f"cloud iteration {shape}", flush=True)
    net = load_network("reduced")
    n = int(np.prod(shape))
    rho = smooth_density(*shape)
    c = CL.Cloud(net, shape, H, rho, gext=(100.0, 0.0, 0.0))
    x0 = workloads.initial_abundances(net, n)
    c.set_abundances(x0)
    st = c.iterate(t_end=1e7, conv_tol=1e-3, precision=lib.MIXED)
    assert st["solved"] == n and st["failed"] == 0 and st["negative_rates"] == 0
    got = c.get(CL.ABUNDANCE)
    # oracle pipeline: its own tracer, columns, field, rates and FP64 solve
    xyz, e, p, ep = oracle_chains(shape)
    slots = CL.shield_slots(net)
    xs = np.where(slots[None, :] >= 0, x0[:, np.maximum(slots, 0)], 0.0)
    Av, cols = O.ray_columns(e, p, ep, rho, xsel=xs)
    G, uv, T = O.uvfield(Av, gext=(100.0, 0.0, 0.0))
    rate = O.Oracle("reduced").rates(T, 20.0, G, Av, cols)
    ref = x0.copy()
    s, ref, _ = O.Oracle("reduced").calculate_abundances(ref, rate, rho, T, mode=O.MODE_GPU)
    assert s == 0
    m = ref[:, :-1] > 1e-20
    err = np.max(np.abs(got[:, :-1][m] - ref[:, :-1][m]) / ref[:, :-1][m])
    print(f"iteration 1 max rel err vs oracle {err:.2e}", flush=True)
    assert err <= 1e-4
    # new columns are those of the new abundances
    xs1 = np.where(slots[None, :] >= 0, got[:, np.maximum(slots, 0)], 0.0)
    _, cols1 = O.ray_columns(e, p, ep, rho, xsel=xs1)
    assert np.allclose(c.get(CL.COLUMNS), cols1, rtol=1e-14, atol=0)
    # second iteration: unconverged points re-solved from their warm start,
    # points not solved keep their abundances bitwise
    st2 = c.iterate(t_end=1e7, conv_tol=1e-3, precision=lib.MIXED)
    got2 = c.get(CL.ABUNDANCE)
    assert st2["failed"] == 0 and st2["solved"] <= n
    moved = np.any(got2 != got, axis=1)
    assert moved.sum() <= st2["solved"]
    # a third iteration with a loose tolerance solves no more points than the second
    st3 = c.iterate(t_end=1e7, conv_tol=1e-1, precision=lib.MIXED)
    assert st3["solved"] <= st2["solved"]
    c.close()


def test_iterations_through_request_loop_equal_direct():
    """The cloud's solves sent through the request loop (device requests,
    graph-launched batches) give the direct launches' abundances bitwise."""
    shape = (5, 4, 4)
    print(f"cloud via server {shape}", flush=True)
    net = load_network("reduced")
    n = int(np.prod(shape))
    rho = smooth_density(*shape)
    x0 = workloads.initial_abundances(net, n)
    out = []
    for via in (False, True):
        c = CL.Cloud(net, shape, H, rho, gext=(100.0, 0.0, 0.0))
        c.set_abundances(x0)
        srv = None
        if via:
            srv = lib.Server(device_bytes=256 << 20, nslots=2, precision=lib.MIXED)
            c.use_server(srv)
        sts = [c.iterate(t_end=1e7, conv_tol=1e-3, precision=lib.MIXED) for _ in range(2)]
        out.append((c.get(CL.ABUNDANCE), [(s["solved"], s["converged"], s["failed"], s["nst"]) for s in sts]))
        c.use_server(None)
        c.close()
        if srv is not None:
            assert srv.graph_stats()[1] >= 2
            srv.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


@pytest.mark.parametrize("order", [lib.ORDER_REFERENCE, lib.ORDER_DISTANCE])
def test_point_cloud_matches_oracle(order):
    """A cloud of arbitrary points (a jittered sphere and a log-spaced 1D slab,
    the shapes of the reference's input files): the device columns, field and
    temperature from the host tracer's evaluation points equal the oracle's
    (in the same candidate order), and one iteration is within the parity bar
    of the FP64 oracle."""
    print(f"point cloud order={order}", flush=True)
    rng = np.random.default_rng(17)
    d = rng.normal(size=(120, 3))
    sphere = d / np.linalg.norm(d, axis=1)[:, None] * rng.uniform(0, 1, (120, 1)) ** (1 / 3) * 0.5
    slab = np.zeros((40, 3))
    slab[1:, 0] = np.logspace(-5, 0, 39)
    net = load_network("reduced")
    for xyz in (sphere, slab):
        n = len(xyz)
        rho = 1e3 * np.exp(0.5 * np.sin(7.0 * xyz[:, 0]) + 0.2 * rng.normal(size=n))
        c = CL.Cloud.from_points(net, xyz, rho, gext=(100.0, 0.0, 0.0), order=order)
        x0 = workloads.initial_abundances(net, n)
        c.set_abundances(x0)
        e, p, ep = O.eval_points(xyz, order=O.ORDER_HEAPSORT if order == lib.ORDER_REFERENCE else O.ORDER_STABLE,
                                 maxpoints=600)
        slots = CL.shield_slots(net)
        xs = np.where(slots[None, :] >= 0, x0[:, np.maximum(slots, 0)], 0.0)
        Av, cols = O.ray_columns(e, p, ep, rho, xsel=xs)
        G, uv, T = O.uvfield(Av, gext=(100.0, 0.0, 0.0))
        assert np.allclose(c.get(CL.AV), Av, rtol=1e-14, atol=0)
        assert np.allclose(c.get(CL.COLUMNS), cols, rtol=1e-14, atol=0)
        assert np.allclose(c.get(CL.TEMPERATURE), T, rtol=1e-13, atol=0)
        st = c.iterate(t_end=1e7, conv_tol=1e-3, precision=lib.MIXED)
        assert st["solved"] == n and st["failed"] == 0
        got = c.get(CL.ABUNDANCE)
        rate = O.Oracle("reduced").rates(T, 20.0, G, Av, cols)
        s_, ref, _ = O.Oracle("reduced").calculate_abundances(x0.copy(), rate, rho, T, mode=O.MODE_GPU)
        assert s_ == 0
        m = ref[:, :-1] > 1e-20
        err = np.max(np.abs(got[:, :-1][m] - ref[:, :-1][m]) / ref[:, :-1][m])
        print(f"  {n} points: max rel err vs oracle {err:.2e}", flush=True)
        assert err <= 1e-4
        c.close()


def test_per_point_zeta_and_pdr_fin(tmp_path):
    """Per-point cosmic-ray rates in the iteration (the cosmic-ray fork's
    per-point zeta) against the oracle evaluated with each point's zeta, and
    the state written as <output>.pdr.fin reads back."""
    from pdrchem import outputs
    shape = (4, 4, 3)
    print(f"cloud per-point zeta {shape}", flush=True)
    net = load_network("reduced")
    n = int(np.prod(shape))
    rho = smooth_density(*shape)
    c = CL.Cloud(net, shape, H, rho, gext=(100.0, 0.0, 0.0))
    x0 = workloads.initial_abundances(net, n)
    c.set_abundances(x0)
    z0 = lib.default_phys().zeta
    zeta = z0 * np.array([1.0, 3.0, 10.0])[np.arange(n) % 3]
    c.set_zeta(zeta)
    st = c.iterate(t_end=1e7, conv_tol=1e-3, precision=lib.MIXED)
    assert st["solved"] == n and st["failed"] == 0
    got = c.get(CL.ABUNDANCE)
    xyz, e, p, ep = oracle_chains(shape)
    slots = CL.shield_slots(net)
    xs = np.where(slots[None, :] >= 0, x0[:, np.maximum(slots, 0)], 0.0)
    Av, cols = O.ray_columns(e, p, ep, rho, xsel=xs)
    G, uv, T = O.uvfield(Av, gext=(100.0, 0.0, 0.0))
    orc = O.Oracle("reduced")
    rate = np.zeros((n, orc.nreac))
    for zf in np.unique(zeta):
        m = zeta == zf
        rate[m] = orc.rates(T[m], 20.0, G[m], Av[m], cols[m], phys=O.default_phys(zeta=float(zf)))
    s_, ref, _ = orc.calculate_abundances(x0.copy(), rate, rho, T, mode=O.MODE_GPU)
    assert s_ == 0
    mk = ref[:, :-1] > 1e-20
    err = np.max(np.abs(got[:, :-1][mk] - ref[:, :-1][mk]) / ref[:, :-1][mk])
    print(f"  max rel err vs oracle {err:.2e}", flush=True)
    assert err <= 1e-4
    path = tmp_path / "cloud.pdr.fin"
    outputs.write_cloud_pdr_fin(str(path), c, xyz)
    back = np.loadtxt(str(path))
    assert back.shape == (n, 9 + net.nspec + 12)
    assert np.allclose(back[:, 9:9 + net.nspec], got, rtol=1e-3, atol=1e-99)
    assert np.allclose(back[:, 1:4], xyz, rtol=1e-3, atol=0)
    c.close()


def test_rejects_bad_arguments():
    net = load_network("reduced")
    with pytest.raises(lib.PdrChemError):
        CL.Cloud(net, (1, 4, 4), H, np.ones(16))
    with pytest.raises(lib.PdrChemError):
        CL.Cloud(net, (4, 4, 4), H, np.ones(64), theta_crit=2.0)
    c = CL.Cloud(net, (3, 3, 3), H, np.full(27, 1e3))
    with pytest.raises(lib.PdrChemError):
        c.iterate()           # abundances not set
    c.close()
