"""Benchmark: grid-point chemistry solves/s on one MI355X.

Default workload (the metric's configuration): the reduced network (33
species / 329 reactions) on the seeded random-point distribution (log10 n_H
~ U[2,6], log10 T ~ U[1,3], A_V ~ U[0,20], log10 G0 ~ U[0,5], zeta =
5e-17 s^-1), 1,000,000 points, initial abundances H = 1, He = 0.1, C+ =
1.4e-4, O = 3e-4, rest 1e-30, integrated to 1e7 yr at rtol 1e-8 / atol 1e-30.
``--network full`` runs the full network (215 species / 2926 reactions) on
the same distribution plus log-uniform N(H2), N(CO) in [1e14, 1e22] cm^-2.

One step = device rate evaluation for every point + reset of the abundances
to the initial state + the batched BDF solve of every point (all inputs
resident in HBM before the timed region).  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "3d-pdr_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 peak (vector and matrix)
FP32_PEAK_TFLOPS = 157.3    # MI355X FP32 peak (matrix)
HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")


def algorithmic_flops(stats, n, n_rhs_flops, n_jac_flops):
    """Work of the reference algorithm (CVODE BDF + Newton + dense LU) for the
    integrator counters of this run:
      W64 = nfe*F_f + nje*F_J + nni*2n^2 + nst*10n      (RHS, Jacobian, solves, Nordsieck)
      WLU = nsetups*(n^2 + 2n^3/3)                      (forming and factoring M)
    F_f, F_J: static op counts of the network's RHS and exact Jacobian."""
    nst, nfe, nje, nsetups, nni = (stats[:, k].astype(np.float64).sum() for k in (0, 1, 2, 3, 4))
    w64 = nfe * n_rhs_flops + nje * n_jac_flops + nni * 2.0 * n * n + nst * 10.0 * n
    wlu = nsetups * (n * n + 2.0 * n ** 3 / 3.0)
    return w64, wlu


def measured_traffic(kernel_name, npts):
    """HBM bytes per launch of the solver kernel from the committed rocprof
    counter summary (bytes per point, copy-kernel calibrated), scaled to npts;
    None when no summary for this kernel exists."""
    path = os.path.join(PROFILES, "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    e = d.get(kernel_name)
    return None if e is None else float(e["bytes_per_point"]) * npts


def network_flops(net):
    terms = sum(len(r.stoich) for r in net.reactions)
    active = sum(1 for r in net.reactions if r.stoich)
    jterms = sum(len(r.stoich) * sum(1 for f in (r.fa, r.fb) if f <= net.neq) for r in net.reactions)
    f_rhs = 2 * active + 2 * terms + 2 * net.neq
    f_jac = 4 * active + 2 * jterms + 2 * net.neq * net.neq
    return f_rhs, f_jac


def per_species_errors(net, got, ref, floor=1e-20):
    """Max relative error per species over the checked points (abundances
    above `floor` in the oracle), largest first; the electron slot is x_e of
    the last RHS evaluation and is reported as well."""
    out = {}
    for i, sp in enumerate(net.species):
        m = ref[:, i] > floor
        if m.any():
            out[sp.name] = float(np.max(np.abs(got[m, i] - ref[m, i]) / ref[m, i]))
    return dict(sorted(out.items(), key=lambda kv: -kv[1]))


def oracle_mode(name):
    import oracle
    return oracle.MODE_GPU if name == "reduced" else oracle.MODE_GPU_WG


def host_cpus():
    """Host CPUs as the box presents them: os.cpu_count(), the CPUs this
    process may run on, the per-GPU CPU share (OMP_NUM_THREADS where the
    environment sets it) and the model name."""
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or aff
    return {"os_cpu_count": os.cpu_count(), "affinity": aff, "share": min(share, aff), "model": model}


def electron_check(name, net, nid, prec, x0, rate, nH, got_charge_sum, idx):
    """The electron slot under the reference's semantics (x_e of the last RHS
    evaluation, calculate_abundances.c:317) on the checked points: a second
    device solve of those points with that policy, against the oracle with
    the same policy, held to the bound the species bar and both sides'
    measured overshoot set (|x_e - q.y| of each side)."""
    import torch
    import oracle
    o = oracle.Oracle(name)
    lib_ = __import__("pdrchem.lib", fromlist=["lib"])
    dev = rate.device
    ab = torch.tensor(x0[idx], device=dev)
    lib_.set_electron_policy(lib_.XE_LAST_RHS)
    lib_.solve_device(nid, ab, rate[idx].contiguous(), nH[idx].contiguous(), precision=prec)
    torch.cuda.synchronize()
    got = ab.cpu().numpy()
    _, ref, _ = o.calculate_abundances(x0[idx].copy(), rate[idx].cpu().numpy(), nH[idx].cpu().numpy(),
                                       nH[idx].cpu().numpy(), mode=oracle_mode(name),
                                       nthreads=host_cpus()["share"], xe_policy=0)
    q = np.array(net.charges(), dtype=np.float64)
    qg, qo = got[:, :-1] @ q, ref[:, :-1] @ q
    diff = np.abs(got[:, -1] - ref[:, -1])
    bound = 1e-4 * qo + np.abs(got[:, -1] - qg) + np.abs(ref[:, -1] - qo)
    same = bool(np.array_equal(got[:, :-1], got_charge_sum[:, :-1]))
    return {"rule": "x_e of the last RHS evaluation (calculate_abundances.c:317), the reference's",
            "max_rel_err_vs_oracle": float(np.max(diff / ref[:, -1])),
            "max_diff_over_measured_bound": float(np.max(diff / bound)),
            "bound": "1e-4*q.y_oracle + |x_e,gpu - q.y_gpu| + |x_e,oracle - q.y_oracle| per point",
            "species_unchanged_by_policy": same, "points": int(len(idx))}


def reference_config_record(name, pts, x0, rates, seconds, threads, mxstep=20000):
    """The reference's own integrator configuration (CVODE's difference-
    quotient Jacobian, plain dense LU: oracle MODE_REFERENCE) on the first
    points of the workload, bounded: mxstep per CVode call (so at most five
    times that many steps per point) and a time budget.  Reported as a
    record, not as the baseline: at atol 1e-30 it fails most points."""
    import oracle
    o = oracle.Oracle(name)
    t0 = time.perf_counter()
    done = ok = 0
    nst = ncf = 0
    flags = []
    while time.perf_counter() - t0 < seconds and done < pts.n:
        sl = slice(done, min(done + threads, pts.n))
        st, _, so = o.calculate_abundances(x0[sl].copy(), rates[sl], pts.n_H[sl], pts.T[sl], mode=oracle.MODE_REFERENCE,
                                           nthreads=threads, mxstep=mxstep)
        ok += int(np.sum(so["flag"] == 0))
        nst += int(so["nst"].sum())
        ncf += int(so["ncfn"].sum())
        flags += [int(f) for f in so["flag"]]
        done = sl.stop
    el = time.perf_counter() - t0
    return {"value": ok / el, "unit": "solves/s", "cores": threads, "kind": "port",
            "attempted": done, "completed": ok, "seconds": el, "mean_steps": nst / max(done, 1),
            "mean_conv_failures": ncf / max(done, 1), "flags": flags[:64],
            "sample": f"first {done} points, CPU oracle in the reference's configuration (difference-quotient "
                      f"Jacobian, plain dense LU), mxstep {mxstep} per CVode call"}


def full_subset_ratios(npts=8192, cpu_seconds=8.0):
    """215-species ratios on a fixed seeded subset of the full-network
    workload (the first npts points of its random-point stream, to 1e7 yr):
    the mixed form served through the request loop (one device request,
    graph launch), the plain FP64 form by a direct launch, and the CPU
    oracle on a bounded sample."""
    import torch
    from pdrchem import lib, workloads
    from pdrchem.netgen import load_network
    dev = torch.device("cuda:0")
    name = "full"
    net = load_network(name)
    nid = lib.network_id(name)
    P = workloads.random_points(npts, columns=True)
    x0 = workloads.initial_abundances(net, npts)
    rate, _ = lib.rates_device(nid, *[torch.tensor(v, device=dev) for v in (P.T, P.Td, P.G, P.Av, P.cols)])
    nH = torch.tensor(P.n_H, device=dev)
    x0d = torch.tensor(x0, device=dev)
    torch.cuda.synchronize()
    # warm-up of both kernels on a few points
    for prec in (lib.MIXED, lib.FP64):
        w = x0d[:64].clone()
        lib.solve_device(nid, w, rate[:64].contiguous(), nH[:64].contiguous(), precision=prec)
    torch.cuda.synchronize()
    srv = lib.Server(device_bytes=2 << 30, nslots=4, precision=lib.MIXED)
    try:
        ab = x0d.clone()
        st = torch.zeros(npts, dtype=torch.int32, device=dev)
        t0 = time.perf_counter()
        srv.submit_device(nid, ab, rate, nH, 1e7, status=st)
        srv.progress(wait=True)
        t_served = time.perf_counter() - t0
        res_served = srv.poll()
    finally:
        srv.close()
    ab64 = x0d.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s64, _ = lib.solve_device(nid, ab64, rate, nH, precision=lib.FP64)
    torch.cuda.synchronize()
    t_fp64 = time.perf_counter() - t0
    cpu = cpu_baseline(name, P, x0, rate.cpu().numpy(), cpu_seconds, host_cpus()["share"])
    v_mixed, v_fp64 = npts / t_served, npts / t_fp64
    return {"subset": f"first {npts} points of the full-network (215 species) workload, to 1e7 yr",
            "mixed_served_solves_s": v_mixed, "fp64_direct_solves_s": v_fp64,
            "mixed_over_fp64": v_mixed / v_fp64, "cpu_solves_s": cpu["value"], "cpu_cores": cpu["cores"],
            "mixed_over_cpu": v_mixed / cpu["value"],
            "failed": int((st != 0).sum().item()) + int((s64 != 0).sum().item()),
            "served_requests": len(res_served)}


def cpu_baseline(name, pts, x0, rates, seconds, threads):
    """Oracle (CPU, same integrator as the GPU path) on a bounded sample of
    the same workload: chunks of points until `seconds` have elapsed."""
    import oracle
    o = oracle.Oracle(name)
    done, t0, chunk = 0, time.perf_counter(), max(threads * 2, 8)
    while time.perf_counter() - t0 < seconds and done < pts.n:
        sl = slice(done, min(done + chunk, pts.n))
        st, _, _ = o.calculate_abundances(x0[sl].copy(), rates[sl], pts.n_H[sl], pts.T[sl],
                                          mode=oracle_mode(name), nthreads=threads)
        done = sl.stop
    el = time.perf_counter() - t0
    how = ("exact Jacobian, error-weight-scaled Gauss-Jordan inverse" if name == "reduced" else
           "exact Jacobian, bordered matrix, static-order head elimination and Gauss-Jordan tail")
    return {"value": done / el, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"first {done} points of the same seeded workload in {el:.1f} s; CPU oracle "
                      f"(C, OpenMP dynamic, -O3) with the GPU path's integrator configuration ({how}); "
                      f"the reference configuration (difference-quotient Jacobian, plain dense LU) "
                      f"stalls on these inputs at atol 1e-30 and is not timed"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--network", choices=["reduced", "full", "mynetwork"], default="reduced")
    ap.add_argument("--npts", type=int, default=None, help="default 1e6 (reduced), 2e4 (others)")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", type=int, default=None,
                    help="points checked against the oracle, evenly spaced over the batch "
                         "(default 1000 reduced, 200 others)")
    ap.add_argument("--electron", choices=["charge-sum", "last-rhs"], default="charge-sum",
                    help="electron slot of the timed solve: charge sum of the returned abundances, or x_e "
                         "of the last RHS evaluation (the reference's); the other is checked on the side")
    ap.add_argument("--no-full-subset", action="store_true",
                    help="skip the 215-species subset ratios of the default run")
    ap.add_argument("--serve", action="store_true",
                    help="serving workload: Poisson stream of requests through the request loop")
    ap.add_argument("--serve-rate", type=float, default=6.0, help="requests per second")
    ap.add_argument("--serve-duration", type=float, default=30.0, help="seconds of arrivals")
    ap.add_argument("--serve-networks", default="reduced,full",
                    help="networks of the requests, chosen with equal odds")
    ap.add_argument("--serve-batch", type=int, default=65536,
                    help="largest batch of the request loop (points)")
    ap.add_argument("--serve-refill", type=int, default=0,
                    help="ring points per network of standing launches (power of two; 0: batches)")
    ap.add_argument("--serve-idle-ms", type=float, default=2.0, help="standing launch idle limit")
    ap.add_argument("--serve-life-ms", type=float, default=250.0, help="standing launch lifetime")
    ap.add_argument("--no-autotune", action="store_true",
                    help="launch with the default shape instead of the autotuner's choice")
    ap.add_argument("--cloud", action="store_true",
                    help="outer chemistry iteration over a lognormal grid cloud (ray-traced columns)")
    ap.add_argument("--cloud-side", type=int, default=64, help="grid side (side^3 points)")
    ap.add_argument("--cloud-iters", type=int, default=8, help="chemistry iterations (CHEMITERATIONS of params.dat)")
    ap.add_argument("--cloud-tol", type=float, default=1e-3, help="convergence tolerance of a point")
    ap.add_argument("--cloud-direct", action="store_true",
                    help="solve the cloud's batches by direct launches instead of through the request loop")
    ap.add_argument("--cloud-thermal", action="store_true",
                    help="thermal balance on (THERMALBALANCE): after the chemistry iterations, the outer "
                         "thermal-balance loop until every point converges or the limits below")
    ap.add_argument("--cloud-thermal-iters", type=int, default=200, help="most thermal-balance iterations")
    ap.add_argument("--cloud-thermal-seconds", type=float, default=300.0,
                    help="no new thermal-balance iteration after this many seconds of the loop")
    ap.add_argument("--cloud-fp64-side", type=int, default=24,
                    help="side of the cloud on which the plain FP64 form is compared (0 = skip)")
    a = ap.parse_args()
    if a.npts is None:
        a.npts = 1_000_000 if a.network == "reduced" else 20_000
    if a.check is None:
        a.check = 1000 if a.network == "reduced" else 200
    if a.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {a.gpus} is not supported "
              f"(grid points are independent: a multi-GPU run would give each GPU its own range of points)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from pdrchem import lib, workloads
    from pdrchem.netgen import load_network

    if a.serve:
        serve(a)
        return
    if a.cloud:
        cloud_bench(a)
        return
    dev = torch.device("cuda:0")
    name = a.network
    net = load_network(name)
    nid = lib.network_id(name)
    prec = lib.MIXED if a.precision == "mixed" else lib.FP64
    pts = workloads.random_points(a.npts, columns=(name != "reduced"))
    x0_np = workloads.initial_abundances(net, a.npts)
    T = torch.tensor(pts.T, device=dev)
    Td = torch.tensor(pts.Td, device=dev)
    G = torch.tensor(pts.G, device=dev)
    Av = torch.tensor(pts.Av, device=dev)
    cols = torch.tensor(pts.cols, device=dev)
    nH = torch.tensor(pts.n_H, device=dev)
    x0 = torch.tensor(x0_np, device=dev)
    ab = torch.empty_like(x0)
    rate = torch.empty((a.npts, net.nreac), dtype=torch.float64, device=dev)
    neg = torch.zeros(1, dtype=torch.int32, device=dev)
    status = torch.zeros(a.npts, dtype=torch.int32, device=dev)
    stats = torch.zeros((a.npts, 8), dtype=torch.int32, device=dev)
    counter = torch.zeros(1, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)
    lib.set_electron_policy(lib.XE_CHARGE_SUM if a.electron == "charge-sum" else lib.XE_LAST_RHS)

    def step(ev=None):
        neg.zero_()
        lib.rates_device(nid, T, Td, G, Av, cols, out=rate, neg=neg, stream=stream)
        ab.copy_(x0)
        counter.zero_()
        if ev:
            ev[0].record(stream)
        lib.solve_device(nid, ab, rate, nH, status=status, stats=stats, counter=counter, precision=prec,
                         stream=stream)
        if ev:
            ev[1].record(stream)

    shape = None
    if not a.no_autotune:
        # launch shape chosen by the autotuner on a sample of the batch (larger
        # batches use the shape of the largest tuned size bucket)
        k = min(a.npts, 40000 if name == "reduced" else 4000)
        lib.rates_device(nid, T[:k], Td[:k], G[:k], Av[:k], cols[:k], out=rate[:k], neg=neg, stream=stream)
        t0 = time.perf_counter()
        r = lib.autotune(nid, x0[:k].clone(), rate[:k], nH[:k], precision=prec)
        shape = {"systems_per_block": r["systems_per_block"], "blocks_per_cu": r["blocks_per_cu"],
                 "chosen_by": f"autotuner on the first {k} points ({time.perf_counter() - t0:.1f} s, untimed)",
                 "applies": lib.tuned_config(nid, a.npts, precision=prec)}
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    times, ktimes = [], []
    t_all0 = time.perf_counter()
    for _ in range(a.steps):
        ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step(ev)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        ktimes.append(ev[0].elapsed_time(ev[1]) / 1e3)
    total = time.perf_counter() - t_all0
    ms = float(np.mean(times) * 1e3)
    st_np = stats.cpu().numpy()
    nfail = int((status != 0).sum().item())
    f_rhs, f_jac = network_flops(net)
    w64, wlu = algorithmic_flops(st_np, net.neq, f_rhs, f_jac)
    kt = float(np.mean(ktimes))
    # roofline of the solver kernel: time the work needs at peak, FP32 for the
    # factorisation in the mixed form, FP64 for everything else
    t_roof = w64 / (FP64_PEAK_TFLOPS * 1e12) + wlu / ((FP32_PEAK_TFLOPS if prec == lib.MIXED else
                                                       FP64_PEAK_TFLOPS) * 1e12)
    achieved = (w64 + wlu) / kt / 1e12
    peak = (w64 + wlu) / t_roof / 1e12
    kname = ("bdf_wave_kernel" if name == "reduced" else "bdf_wg_kernel") + f"<{name},{a.precision}>"
    wl = {"reduced": "reduced network (33 species/329 reactions)",
          "full": "full network (215 species/2926 reactions), N(H2), N(CO) ~ logU[1e14,1e22]",
          "mynetwork": "MYNETWORK (58 species/634 reactions)"}[name]
    res = {
        "metric": "grid-point chemistry solves/s at <=1e-4 rel err vs FP64 CVODE; p50/p99 batch ms",
        "value": a.npts / float(np.mean(times)),
        "unit": "solves/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms,
        "p50_batch_ms": float(np.percentile(np.array(times) * 1e3, 50)),
        "p99_batch_ms": float(np.percentile(np.array(times) * 1e3, 99)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == lib.FP64 else "f64+f32-matrix",
        "data": "synthetic (seeded counter-based RNG)",
        "config": {"workload": wl + ", random points log10 nH~U[2,6], log10 T~U[1,3], Av~U[0,20], "
                               "log10 G0~U[0,5], to 1e7 yr",
                   "points": a.npts, "precision": a.precision, "rtol": 1e-8, "atol": 1e-30},
        "failed_points": nfail,
        "electron_policy": ("charge sum of the returned abundances (x_e at t_end)" if a.electron == "charge-sum"
                            else "x_e of the last RHS evaluation (calculate_abundances.c:317)"),
        "launch_shape": shape,
        "mean_steps_per_point": float(st_np[:, 0].mean()),
        "solver_kernel_ms": kt * 1e3,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": measured_traffic(kname, a.npts),
                     "kernel": kname,
                     "note": "work W of CVODE's algorithm (dense LU) for this run's integrator counters; "
                             "peak blends FP64 78.6 and (mixed form, factorisation) FP32 157.3 TFLOP/s"},
    }
    cpus = host_cpus()
    res["host_cpus"] = cpus
    if a.check > 0:
        import oracle
        o = oracle.Oracle(name)
        k = min(a.check, a.npts)
        idx = np.unique(np.linspace(0, a.npts - 1, k).astype(np.int64))
        got = ab[idx].cpu().numpy()
        xp = 1 if a.electron == "charge-sum" else 0
        _, ref, _ = o.calculate_abundances(x0_np[idx].copy(), rate[idx].cpu().numpy(), pts.n_H[idx], pts.T[idx],
                                           mode=oracle_mode(name), nthreads=cpus["share"], xe_policy=xp)
        slots = slice(None) if xp == 1 else slice(0, -1)
        g, r_ = got[:, slots], ref[:, slots]
        m = r_ > 1e-20
        res["max_rel_err_vs_oracle"] = float(np.max(np.abs(g[m] - r_[m]) / r_[m]))
        res["max_rel_err_slots"] = ("all species including the electron (charge-sum policy)" if xp == 1 else
                                    "all species except the electron slot (reported under electron_check)")
        res["checked_points"] = int(len(idx))
        res["checked_points_rule"] = f"{len(idx)} points evenly spaced over the {a.npts}-point batch"
        res["points_over_bar"] = int(np.sum(np.any(np.where(m, np.abs(g - r_) / np.where(m, r_, 1.0), 0.0) > 1e-4,
                                                   axis=1)))
        res["per_species_max_rel_err"] = per_species_errors(net, got, ref)
        if xp == 1:
            sel = idx[: min(len(idx), 256)]
            res["electron_check"] = electron_check(name, net, nid, prec, x0_np, rate, nH, got[: len(sel)], sel)
            lib.set_electron_policy(lib.XE_CHARGE_SUM)
    if not a.no_cpu:
        thr = cpus["share"]
        k = min(a.npts, 200000)
        res["cpu_baseline"] = cpu_baseline(name, pts.subset(np.arange(k)), x0_np[:k], rate[:k].cpu().numpy(),
                                           a.cpu_seconds, thr)
        res["cpu_baseline"]["cores_note"] = (f"{thr} threads = the CPU share of this GPU (OMP_NUM_THREADS / "
                                             f"affinity); os.cpu_count() {cpus['os_cpu_count']}")
        if name == "reduced":
            res["cpu_baseline_reference"] = reference_config_record(name, pts.subset(np.arange(k)), x0_np[:k],
                                                                    rate[:k].cpu().numpy(), 10.0, thr)
    if name == "reduced" and not a.no_full_subset and prec == lib.MIXED:
        res["full_network_subset"] = full_subset_ratios()
    lib.set_electron_policy(lib.XE_LAST_RHS)
    print(json.dumps(res))


def serve(a):
    """Serving workload: requests of 1..4096 points (log-uniform) of the
    networks in --serve-networks (equal odds), end times log-uniform in
    [1e3, 1e7] yr, Poisson arrivals, submitted to the request loop as they
    arrive; latency is submit-to-result, throughput is points per second
    from the first arrival to the last result.  Inputs (device-evaluated
    rates) are prepared before the stream starts."""
    import torch
    from pdrchem import lib, workloads
    from pdrchem.netgen import load_network
    dev = torch.device("cuda:0")
    prec = lib.MIXED if a.precision == "mixed" else lib.FP64
    nets = tuple(a.serve_networks.split(","))
    R = workloads.request_stream(duration=a.serve_duration, rate=a.serve_rate, networks=nets)
    data = {}
    for name in nets:
        tot = sum(r.npts for r in R if r.network == name)
        net = load_network(name)
        nid = lib.network_id(name)
        P = workloads.random_points(max(tot, 1), columns=(name != "reduced"))
        rate, neg = lib.rates_device(nid, *[torch.tensor(v, device=dev) for v in (P.T, P.Td, P.G, P.Av, P.cols)])
        torch.cuda.synchronize()
        data[name] = dict(nid=nid, rate=rate.cpu().numpy(), nH=P.n_H.copy(),
                          ab=workloads.initial_abundances(net, max(tot, 1)))
    lib.set_globals(rtol=1e-8, atol=1e-30)
    tuned = {}
    if not a.no_autotune:
        # launch shapes for the request sizes of the stream, tuned before it starts
        for name in nets:
            d = data[name]
            for n in (64, 1024, 4096):
                n = min(n, d["ab"].shape[0])
                r = lib.autotune(d["nid"], torch.tensor(d["ab"][:n], device=dev), torch.tensor(d["rate"][:n], device=dev),
                                 torch.tensor(d["nH"][:n], device=dev), precision=prec)
                tuned[f"{name}:{n}"] = (r["systems_per_block"], r["blocks_per_cu"])
    srv = lib.Server(device_bytes=4 << 30, nslots=4, precision=prec)
    srv.set_max_batch(a.serve_batch)
    if a.serve_refill:
        srv.set_refill(a.serve_refill, idle_ms=a.serve_idle_ms, life_ms=a.serve_life_ms)
    lat, status = {}, {}
    size = {}
    backlog = []            # (seconds, points submitted and not yet done)
    t0 = time.perf_counter()
    i = 0
    open_pts = 0
    t_report = 0.0
    while i < len(R) or len(lat) < len(R):
        now = time.perf_counter() - t0
        while i < len(R) and R[i].arrival <= now:
            r, d = R[i], data[R[i].network]
            sl = slice(r.first, r.first + r.npts)
            tk = srv.submit(d["nid"], d["ab"][sl], d["rate"][sl], d["nH"][sl], r.t_end)
            size[tk] = (r.npts, r.network)
            open_pts += r.npts
            i += 1
        srv.progress()
        for tk, st, ms in srv.poll():
            lat[tk], status[tk] = ms, st
            open_pts -= size[tk][0]
        backlog.append((now, open_pts))
        if now - t_report > 30.0:
            t_report = now
            print(f"serving: {now:.0f} s, {i} of {len(R)} submitted, {len(lat)} done, backlog {open_pts} points",
                  file=sys.stderr, flush=True)
        if i < len(R):
            time.sleep(max(0.0, min(0.002, R[i].arrival - (time.perf_counter() - t0))))
        else:
            time.sleep(0.001)
    wall = time.perf_counter() - t0
    refill = srv.refill_stats()
    srv.close()
    L = np.array([lat[k] for k in sorted(lat)])
    npts = sum(r.npts for r in R)
    B = np.array(backlog)
    t_last = R[-1].arrival
    during = B[B[:, 0] <= t_last]
    # queue growth over the arrival window: least-squares slope of the backlog
    growth = float(np.polyfit(during[:, 0], during[:, 1], 1)[0]) if len(during) > 2 else 0.0
    res = {
        "metric": "grid-point chemistry solves/s at <=1e-4 rel err vs FP64 CVODE; p50/p99 batch ms",
        "value": npts / wall, "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": wall * 1e3, "p50_batch_ms": float(np.percentile(L, 50)),
        "p99_batch_ms": float(np.percentile(L, 99)), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if prec == lib.FP64 else "f64+f32-matrix",
        "data": "synthetic (seeded counter-based RNG)",
        "config": {"workload": f"serving: Poisson {a.serve_rate} req/s for {a.serve_duration} s, 1-4096 points "
                               f"per request (log-uniform), networks {'/'.join(nets)} with equal odds, "
                               f"t_end logU[1e3,1e7] yr",
                   "requests": len(R), "points": npts, "precision": a.precision, "slots": 4,
                   "max_batch": a.serve_batch},
        "failed_requests": int(sum(1 for v in status.values() if v != 0)),
        "launch_shapes": tuned or None,
        "mode": (f"standing launches (ring {a.serve_refill} points per network, idle {a.serve_idle_ms} ms, "
                 f"life {a.serve_life_ms} ms)" if a.serve_refill else "batches"),
        "standing_launches": refill[1] if a.serve_refill else None,
        "backlog_points_at_last_arrival": int(during[-1, 1]) if len(during) else 0,
        "backlog_points_max": int(B[:, 1].max()),
        "backlog_growth_points_per_s": growth,
        "arrival_points_per_s": npts / max(t_last, 1e-9),
        "per_network": {nm: {"requests": len(v), "p50_ms": float(np.percentile(v, 50)),
                             "p99_ms": float(np.percentile(v, 99))}
                        for nm in nets for v in [[lat[k] for k in lat if size[k][1] == nm]] if v},
    }
    print(json.dumps(res))


def cloud_bench(a):
    """Outer chemistry iteration over a synthetic cloud (the chemistry
    iterations of 3DPDR.F90:474-520 with the ray-traced Av and columns of
    eval_points.F90 / 3DPDR.F90:1636-1844): a side^3 grid of 1/16 pc cells,
    lognormal n_H (mean 1e3 cm^-3, sigma_ln 1, correlation 4 cells), G0 = 100
    Draine on the -x face (UNI field along +x), HEALPix level 0 (12 rays),
    theta_crit 1.3, T from GUESS_TEMP, T_d = 20 K, full network by default.
    Every iteration computes rates from the current columns, solves the
    points that are not converged (warm-started from their abundances, t = 0
    to 1e7 yr), writes back and recomputes the columns; state stays on the
    device.  value = points solved / wall time of the iterations; the ray
    tracer (set-up) is timed separately."""
    from pdrchem import cloud as CL, lib, workloads
    from pdrchem.netgen import load_network
    import oracle
    name = a.network if a.network != "reduced" or "--network" in sys.argv else "full"
    net = load_network(name)
    prec = lib.MIXED if a.precision == "mixed" else lib.FP64
    side = a.cloud_side
    lib.set_globals(rtol=1e-8, atol=1e-30)

    def make(sd):
        rho = workloads.cloud_density(sd)
        t0 = time.perf_counter()
        c = CL.Cloud(net, (sd, sd, sd), 0.0625, rho, gext=(100.0, 0.0, 0.0))
        t_setup = time.perf_counter() - t0
        x0 = workloads.initial_abundances(net, sd ** 3)
        c.set_abundances(x0)
        return c, rho, x0, t_setup

    c, rho, x0, t_setup = make(side)
    if a.cloud_thermal:
        # coolants, temperature brackets and dust temperatures from the field
        # (THERMALBALANCE, GUESS_TEMP, DUST) before the first chemistry
        # iteration, as the reference's main program sets them up
        c.thermal_init(params=lib.default_thermal_params(pseudo_ray=-1))
    srv = None
    if not a.cloud_direct:
        # the shrinking, uneven batches of the iterations go through the request
        # loop as device requests (state stays in HBM; graph-launched batches)
        srv = lib.Server(device_bytes=8 << 30, nslots=4, precision=prec)
        c.use_server(srv)
    T = c.get(CL.TEMPERATURE)
    G = c.get(CL.G)
    Av = c.get(CL.AV)
    cols0 = c.get(CL.COLUMNS)
    Td = c.get(CL.DUST_TEMPERATURE) if a.cloud_thermal else np.full(side ** 3, 20.0)
    n_ep = int(lib.load().pdrchem_cloud_num_evaluation_points(c._h))
    iters, wall, solved = [], 0.0, 0
    tot = dict(nst=0, nfe=0, nje=0, nsetups=0, nni=0)
    ab1 = None
    for it in range(a.cloud_iters):
        t0 = time.perf_counter()
        st = c.iterate(t_end=1e7, conv_tol=a.cloud_tol, precision=prec)
        dt = time.perf_counter() - t0
        wall += dt
        solved += st["solved"]
        for k in tot:
            tot[k] += st[k]
        iters.append({"iteration": it + 1, "solved": st["solved"], "converged": st["converged"],
                      "failed": st["failed"], "wall_ms": dt * 1e3, "solve_ms": st["solve_ms"],
                      "rates_ms": st["rates_ms"]})
        print(f"cloud iteration {it + 1}: solved {st['solved']} converged {st['converged']} "
              f"failed {st['failed']} in {dt:.2f} s", file=sys.stderr, flush=True)
        if it == 0:
            ab1 = c.get(CL.ABUNDANCE)
    thermal = None
    if a.cloud_thermal:
        # the outer thermal-balance loop (3DPDR.F90:590-1241): level
        # populations, cooling, heating and the temperature update every
        # iteration; chemistry of the thermally unconverged points when the
        # populations have converged; converged points drop out
        th_iters, th_wall, th_solves = [], 0.0, 0
        t_loop = time.perf_counter()
        for it in range(a.cloud_thermal_iters):
            if time.perf_counter() - t_loop > a.cloud_thermal_seconds:
                break
            t0 = time.perf_counter()
            st = c.thermal_iterate(t_end=1e7, precision=prec)
            dt = time.perf_counter() - t0
            th_wall += dt
            th_solves += st["chemistry_solves"]
            th_iters.append({"iteration": st["iteration"], "lvg": st["lvg"], "levels_converged": st["level_converged"],
                             "thermally_converged": st["thermally_converged"], "updated": st["thermal_points"],
                             "chemistry_solves": st["chemistry_solves"], "failed": st["failed"],
                             "wall_ms": dt * 1e3})
            if st["chemistry_solves"] or it % 10 == 0 or st["done"]:
                print(f"thermal iteration {st['iteration']}: levels {st['level_converged']}, thermally converged "
                      f"{st['thermally_converged']} of {side ** 3}, chemistry solves {st['chemistry_solves']} "
                      f"in {dt:.2f} s", file=sys.stderr, flush=True)
            if st["done"]:
                break
        Tf = c.get(CL.TEMPERATURE)
        thermal = {"iterations": len(th_iters), "done": bool(th_iters and th_iters[-1]["thermally_converged"] ==
                                                          side ** 3) or bool(st["done"]),
                   "thermally_converged": th_iters[-1]["thermally_converged"] if th_iters else 0,
                   "chemistry_solves": th_solves, "wall_s": th_wall,
                   "chemistry_solves_per_s": th_solves / th_wall if th_wall else 0.0,
                   "iterations_per_s": len(th_iters) / th_wall if th_wall else 0.0,
                   "T_gas_percentiles_K": [float(np.percentile(Tf, q)) for q in (1, 50, 99)],
                   "chemistry_rounds": [i for i in th_iters if i["chemistry_solves"]]}
    f_rhs, f_jac = network_flops(net)
    stt = np.array([[tot["nst"], tot["nfe"], tot["nje"], tot["nsetups"], tot["nni"], 0, 0, 0]], dtype=np.float64)
    w64, wlu = algorithmic_flops(stt, net.neq, f_rhs, f_jac)
    ksec = sum(i["solve_ms"] for i in iters) / 1e3
    t_roof = w64 / (FP64_PEAK_TFLOPS * 1e12) + wlu / ((FP32_PEAK_TFLOPS if prec == lib.MIXED else
                                                       FP64_PEAK_TFLOPS) * 1e12)
    achieved = (w64 + wlu) / ksec / 1e12
    peak = (w64 + wlu) / t_roof / 1e12
    kname = ("bdf_wave_kernel" if name == "reduced" else "bdf_wg_kernel") + f"<{name},{a.precision}>"
    res = {
        "metric": "grid-point chemistry solves/s at <=1e-4 rel err vs FP64 CVODE; p50/p99 batch ms",
        "value": solved / wall, "unit": "solves/s", "n_gpus": 1, "steps": a.cloud_iters, "warmup": 0,
        "ms_per_step": wall * 1e3 / max(len(iters), 1),
        "p50_batch_ms": float(np.percentile([i["wall_ms"] for i in iters], 50)),
        "p99_batch_ms": float(np.percentile([i["wall_ms"] for i in iters], 99)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == lib.FP64 else "f64+f32-matrix",
        "data": "synthetic (seeded counter-based RNG)",
        "config": {"workload": f"outer chemistry iteration over a {side}^3 lognormal cloud (n_H mean 1e3, "
                               f"sigma_ln 1), 1/16 pc cells, G0=100 on the -x face, 12 HEALPix rays, "
                               f"{name} network, warm starts, converged points skipped (tol {a.cloud_tol})",
                   "points": side ** 3, "iterations": a.cloud_iters, "precision": a.precision,
                   "rtol": 1e-8, "atol": 1e-30,
                   "solves": "direct launches" if srv is None else "request loop (device requests, 4 slots)"},
        "iterations": iters,
        "thermal_balance": thermal,
        "raytrace_setup_s": t_setup, "evaluation_points": n_ep,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": measured_traffic(kname, solved // max(len(iters), 1)),
                     "kernel": kname,
                     "note": "work W of CVODE's algorithm (dense LU) for the counters of every solve of the run"},
    }
    # iteration 1 checked point by point against the FP64 oracle fed with the
    # same inputs (device T, G, Av, columns; oracle rates)
    k = min(a.check, 256, side ** 3) if a.check > 0 else 0
    thr = host_cpus()["share"]
    if k:
        sel = np.linspace(0, side ** 3 - 1, k).astype(int)
        o = oracle.Oracle(name)
        rate = o.rates(T[sel], Td[sel], G[sel], Av[sel], cols0[sel])
        _, ref, _ = o.calculate_abundances(x0[sel].copy(), rate, rho[sel], T[sel], mode=oracle_mode(name),
                                           nthreads=thr)
        got = ab1[sel]
        m = ref[:, :-1] > 1e-20
        res["max_rel_err_vs_oracle"] = float(np.max(np.abs(got[:, :-1][m] - ref[:, :-1][m]) / ref[:, :-1][m]))
        res["checked_points"] = k
        res["per_species_max_rel_err"] = per_species_errors(net, got, ref)
    if not a.no_cpu:
        o = oracle.Oracle(name)
        order = np.random.default_rng(7).permutation(side ** 3)
        done, t0, chunk = 0, time.perf_counter(), thr * 2
        while time.perf_counter() - t0 < a.cpu_seconds and done < side ** 3:
            sel = order[done:done + chunk]
            rate = o.rates(T[sel], Td[sel], G[sel], Av[sel], cols0[sel])
            o.calculate_abundances(x0[sel].copy(), rate, rho[sel], T[sel], mode=oracle_mode(name), nthreads=thr)
            done += len(sel)
        el = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": done / el, "unit": "solves/s", "cores": thr, "kind": "port",
                               "sample": f"{done} random points of the cloud's first iteration (oracle rates and "
                                         f"solve) in {el:.1f} s on {thr} threads"}
    if a.cloud_fp64_side > 0 and prec == lib.MIXED:
        # plain FP64 form against the mixed form on the same (smaller) cloud, first iteration
        sd = a.cloud_fp64_side
        out = {}
        for pname, pv in (("mixed", lib.MIXED), ("fp64", lib.FP64)):
            cc, _, _, _ = make(sd)
            t0 = time.perf_counter()
            st = cc.iterate(t_end=1e7, conv_tol=a.cloud_tol, precision=pv)
            out[pname] = st["solved"] / (time.perf_counter() - t0)
            cc.close()
        res["fp64_comparison"] = {"points": sd ** 3, "iteration": 1, "mixed_solves_per_s": out["mixed"],
                                  "fp64_solves_per_s": out["fp64"], "ratio": out["mixed"] / out["fp64"]}
    c.use_server(None)
    c.close()
    if srv is not None:
        res["request_loop_graph_launches"] = srv.graph_stats()[1]
        srv.close()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
