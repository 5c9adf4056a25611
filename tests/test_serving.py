"""Request loop with continuous batching (pdrchem_server_* in libpdrchem.so).

CPU: the seeded request stream of the serving workload (Poisson arrivals,
log-uniform sizes and end times) and the library's exports.
GPU: requests of two networks and many end times submitted while others run
are batched together on several streams, or taken point by point by
standing launches; every request's result must be bit-identical to the same points solved alone with the same end time (each
point is integrated by its own wavefront / workgroup with a fixed order of
operations, whatever else shares the launch), and the bookkeeping (tickets,
statuses, latencies) must account for every request.
"""
import numpy as np
import pytest

from pdrchem import workloads


def test_request_stream_properties():
    R = workloads.request_stream(duration=30.0, rate=200.0)
    assert R == workloads.request_stream(duration=30.0, rate=200.0)  # seeded
    arr = np.array([r.arrival for r in R])
    assert np.all(np.diff(arr) >= 0) and arr[-1] < 30.0
    assert abs(len(R) / 30.0 - 200.0) < 200.0 * 0.1              # Poisson count, 6000 expected
    n = np.array([r.npts for r in R])
    assert n.min() >= 1 and n.max() <= 4096
    assert 400 < n.mean() < 600                                   # E = 4095/ln 4096 ~ 492
    te = np.array([r.t_end for r in R])
    assert te.min() >= 1e3 and te.max() <= 1e7
    frac_full = np.mean([r.network == "full" for r in R])
    assert 0.45 < frac_full < 0.55
    # each network's requests take consecutive, disjoint point ranges
    for nm in ("reduced", "full"):
        sub = [r for r in R if r.network == nm]
        assert all(b.first == a.first + a.npts for a, b in zip(sub, sub[1:]))


def _requests(orc_by_net, nets_sizes, seed):
    from pdrchem.netgen import load_network
    rng = np.random.default_rng(seed)
    out = []
    for k, (name, n) in enumerate(nets_sizes):
        net = load_network(name)
        P = workloads.random_points(n, seed=seed + k, columns=(name != "reduced"))
        x0 = workloads.initial_abundances(net, n)
        r = orc_by_net[name].rates(P.T, P.Td, P.G, P.Av, P.cols)
        t_end = float(10.0 ** rng.uniform(3, 7))
        out.append((name, x0, np.ascontiguousarray(r), P.n_H.copy(), t_end))
    return out


@pytest.mark.gpu
def test_server_mix_equals_alone():
    import torch
    import oracle
    from pdrchem import lib
    orc = {nm: oracle.Oracle(nm) for nm in ("reduced", "full")}
    sizes = [("reduced", 37), ("full", 3), ("reduced", 200), ("reduced", 1), ("full", 5),
             ("reduced", 64), ("full", 2), ("reduced", 9)]
    reqs = _requests(orc, sizes, seed=42)
    lib.set_globals(rtol=1e-8, atol=1e-30)
    srv = lib.Server(device_bytes=256 << 20, nslots=3, precision=lib.MIXED)
    abs_mix, tickets = [], []
    for name, x0, r, nH, te in reqs:
        ab = x0.copy()
        abs_mix.append(ab)
        tickets.append(srv.submit(lib.network_id(name), ab, r, nH, te))
        srv.progress()
    srv.progress(wait=True)
    done = srv.poll()
    graphs, launches, _ = srv.graph_stats()
    srv.close()
    assert graphs and launches >= 1
    assert sorted(t for t, _, _ in done) == sorted(tickets)
    assert all(st == 0 for _, st, _ in done)
    assert all(lat > 0 for _, _, lat in done)
    # the same requests, each alone through the device entry point
    dev = torch.device("cuda:0")
    for (name, x0, r, nH, te), ab_mix in zip(reqs, abs_mix):
        ab = torch.tensor(x0, device=dev)
        status, _ = lib.solve_device(lib.network_id(name), ab, torch.tensor(r, device=dev),
                                     torch.tensor(nH, device=dev), t_end=te, precision=lib.MIXED)
        torch.cuda.synchronize()
        assert int(status.abs().sum().item()) == 0
        alone = ab.cpu().numpy()
        assert np.array_equal(alone, ab_mix), (name, te, float(np.max(np.abs(alone - ab_mix))))


@pytest.mark.gpu
def test_server_graph_replays_equal_stream_launches(monkeypatch):
    """One slot, so its executable graph is updated and replayed batch after
    batch (different sizes, networks and end times, with host work and
    synchronisation between replays); every replay's results equal those of
    plain stream launches bitwise."""
    import oracle
    from pdrchem import lib
    orc = {nm: oracle.Oracle(nm) for nm in ("reduced", "full")}
    sizes = [("reduced", 50), ("reduced", 50), ("reduced", 13), ("full", 4), ("full", 4), ("reduced", 50)]
    reqs = _requests(orc, sizes, seed=7)
    lib.set_globals(rtol=1e-8, atol=1e-30)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("PDRCHEM_SERVER_GRAPHS", mode)
        srv = lib.Server(device_bytes=256 << 20, nslots=1, precision=lib.MIXED)
        res = []
        for name, x0, r, nH, te in reqs:
            ab = x0.copy()
            srv.submit(lib.network_id(name), ab, r, nH, te)
            srv.progress(wait=True)          # one batch per request: a replay each
            assert all(st == 0 for _, st, _ in srv.poll())
            res.append(ab)
        g, launches, inst = srv.graph_stats()
        srv.close()
        assert g == (mode == "1")
        if g:
            assert launches == len(reqs) and 1 <= inst <= len(reqs)
        out[mode] = res
    for a, b in zip(out["0"], out["1"]):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_device_requests_match_host_requests():
    """Requests whose arrays live on the device (staged device to device,
    abundances written back in place, per-point statuses and counters
    returned) give the host requests' results bitwise, also when both kinds
    are in flight in one server."""
    import torch
    import oracle
    from pdrchem import lib
    orc = {nm: oracle.Oracle(nm) for nm in ("reduced", "full")}
    sizes = [("reduced", 40), ("full", 3), ("reduced", 120), ("reduced", 7), ("full", 2)]
    reqs = _requests(orc, sizes, seed=11)
    lib.set_globals(rtol=1e-8, atol=1e-30)
    dev = torch.device("cuda:0")
    srv = lib.Server(device_bytes=256 << 20, nslots=2, precision=lib.MIXED)
    host_out, dev_out = [], []
    for name, x0, r, nH, te in reqs:
        ab = x0.copy()
        host_out.append(ab)
        srv.submit(lib.network_id(name), ab, r, nH, te)
        d = dict(ab=torch.tensor(x0, device=dev), r=torch.tensor(r, device=dev), nH=torch.tensor(nH, device=dev),
                 st=torch.full((len(nH),), 99, dtype=torch.int32, device=dev),
                 stats=torch.zeros((len(nH), 8), dtype=torch.int32, device=dev))
        dev_out.append(d)
        srv.submit_device(lib.network_id(name), d["ab"], d["r"], d["nH"], te, status=d["st"], stats=d["stats"])
        srv.progress()
    srv.progress(wait=True)
    done = srv.poll()
    srv.close()
    torch.cuda.synchronize()
    assert len(done) == 2 * len(reqs) and all(st == 0 for _, st, _ in done)
    for h, d in zip(host_out, dev_out):
        assert np.array_equal(h, d["ab"].cpu().numpy())
        assert int(d["st"].abs().sum().item()) == 0
        assert int(d["stats"][:, 0].min().item()) > 0          # nst of every point


@pytest.mark.gpu
def test_server_rejects_oversized():
    from pdrchem import lib
    srv = lib.Server(device_bytes=1 << 20, nslots=1)
    n = 4096
    ab = np.zeros((n, 215))
    with pytest.raises(lib.PdrChemError):
        srv.submit(lib.network_id("full"), ab, np.zeros((n, 2926)), np.ones(n), 1e3)
    srv.close()


def _drain(srv, n, seconds=120.0):
    """progress/poll until n requests are done (bounded: fails rather than
    waits for ever if results stop coming)"""
    import time
    done, t0 = [], time.perf_counter()
    while len(done) < n:
        assert time.perf_counter() - t0 < seconds, f"{len(done)} of {n} requests done after {seconds} s"
        srv.progress()
        done += srv.poll()
        time.sleep(0.001)
    return done


def _alone(name, x0, r, nH, te, prec):
    import torch
    from pdrchem import lib
    dev = torch.device("cuda:0")
    ab = torch.tensor(x0, device=dev)
    status, _ = lib.solve_device(lib.network_id(name), ab, torch.tensor(r, device=dev), torch.tensor(nH, device=dev),
                                 t_end=te, precision=prec)
    torch.cuda.synchronize()
    return ab.cpu().numpy(), status.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("prec_name", ["mixed", "fp64"])
def test_refill_requests_equal_alone(prec_name):
    """Standing launches (pdrchem_server_set_refill): requests of two networks
    and many end times, more points than a ring holds (so requests wrap round
    the ring's end and wait for space), submitted in two waves, the second
    while the first runs.  Every request's result is bitwise that of the same
    points solved alone, every ticket comes back once with status 0, and no
    point stays held."""
    import time
    import oracle
    from pdrchem import lib
    prec = lib.MIXED if prec_name == "mixed" else lib.FP64
    orc = {nm: oracle.Oracle(nm) for nm in ("reduced", "full")}
    sizes = [("reduced", 37), ("full", 3), ("reduced", 200), ("reduced", 1), ("full", 5), ("reduced", 64),
             ("full", 2), ("reduced", 9), ("reduced", 150), ("full", 60), ("reduced", 100), ("full", 70)]
    reqs = _requests(orc, sizes, seed=5)
    lib.set_globals(rtol=1e-8, atol=1e-30)
    srv = lib.Server(device_bytes=64 << 20, nslots=1, precision=prec)
    srv.set_refill(256, idle_ms=5.0, life_ms=200.0)
    print(f"refill {prec_name}: ring 256, {len(reqs)} requests", flush=True)
    outs, tickets = [], []
    for k, (name, x0, r, nH, te) in enumerate(reqs):
        if k == len(reqs) // 2:
            srv.progress()
            time.sleep(0.05)                       # the first wave is running
        ab = x0.copy()
        outs.append(ab)
        tickets.append(srv.submit(lib.network_id(name), ab, r, nH, te))
        srv.progress()
    done = _drain(srv, len(reqs))
    ring, launches, held = srv.refill_stats()
    srv.close()
    print(f"  {launches} standing launches", flush=True)
    assert ring == 256 and held == 0 and launches >= 2
    assert sorted(t for t, _, _ in done) == sorted(tickets)
    assert all(st == 0 for _, st, _ in done) and all(lat > 0 for _, _, lat in done)
    for (name, x0, r, nH, te), got in zip(reqs, outs):
        alone, st = _alone(name, x0, r, nH, te, prec)
        assert not st.any()
        assert np.array_equal(alone, got), (name, te, float(np.max(np.abs(alone - got))))


@pytest.mark.gpu
def test_refill_failure_statuses_equal_alone():
    """With a step limit that stops some points (pdrchem_set_mxstep), a
    standing launch reports each request's first failing status and the
    partial abundances exactly as the same points solved alone."""
    import oracle
    from pdrchem import lib
    reqs = _requests({"reduced": oracle.Oracle("reduced")}, [("reduced", 40), ("reduced", 25)], seed=9)
    lib.set_globals(rtol=1e-8, atol=1e-30)
    old = lib.mxstep()
    lib.set_mxstep(60)
    try:
        srv = lib.Server(device_bytes=64 << 20, nslots=1, precision=lib.MIXED)
        srv.set_refill(64)
        outs, tickets = [], []
        for name, x0, r, nH, te in reqs:
            ab = x0.copy()
            outs.append(ab)
            tickets.append(srv.submit(lib.network_id(name), ab, r, nH, 1e7))
        done = {t: st for t, st, _ in _drain(srv, len(reqs))}
        srv.close()
        for (name, x0, r, nH, te), got, tk in zip(reqs, outs, tickets):
            alone, st = _alone(name, x0, r, nH, 1e7, lib.MIXED)
            first = int(st[np.nonzero(st)[0][0]]) if st.any() else 0
            assert first != 0 and done[tk] == first
            assert np.array_equal(alone, got)
    finally:
        lib.set_mxstep(old)


@pytest.mark.gpu
def test_refill_rejects_bad_settings():
    from pdrchem import lib
    srv = lib.Server(device_bytes=1 << 20, nslots=1)
    for ring in (100, 32, -1, 1 << 25):
        with pytest.raises(lib.PdrChemError):
            srv.set_refill(ring)
    srv.set_refill(64)
    with pytest.raises(lib.PdrChemError):
        srv.submit(lib.network_id("reduced"), np.zeros((65, 33)), np.zeros((65, 329)), np.ones(65), 1e3)
    assert srv.refill_stats() == (64, 0, 0)
    srv.set_refill(0)
    srv.close()
