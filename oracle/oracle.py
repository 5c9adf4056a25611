"""ctypes binding of the CPU oracle (liboracle.so) and of the reference rate
code built into ``oracle/_ref`` (when present).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product path never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
NET_DIR = os.path.join(ROOT, "3d-pdr_amd", "networks")
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN_DIR = os.path.join(HERE, "_ref")

SECONDS_IN_YEAR = 3.1556926e7

# integrator method bits (pdr_oracle.h)
MODE_REFERENCE = 0          # CVODE difference-quotient Jacobian, plain dense LU
MODE_EXACT_JAC = 1
MODE_SCALED = 2
MODE_F32 = 16
MODE_GJ = 64
MODE_GPU = 1 | 2 | MODE_GJ       # exact J, weight-scaled matrix, Gauss-Jordan inverse
MODE_GPU_MIXED = MODE_GPU | MODE_F32
MODE_STATIC = 128
MODE_HYBRID = 256
MODE_GPU_WG = 1 | 2 | MODE_HYBRID  # bordered matrix, static head + Gauss-Jordan tail
MODE_GPU_WG_MIXED = MODE_GPU_WG | MODE_F32

def MODE_REFINE(k):
    return k << 2

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Phys(C.Structure):
    _fields_ = [("zeta", C.c_double), ("v_turb", C.c_double), ("omega", C.c_double),
                ("grain_radius", C.c_double), ("metallicity", C.c_double), ("g2d", C.c_double)]


class Opts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("mxstep", C.c_long),
                ("mode", C.c_int), ("maxord", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("nst", C.c_long), ("nfe", C.c_long), ("nfeDQ", C.c_long), ("nsetups", C.c_long),
                ("nje", C.c_long), ("nni", C.c_long), ("ncfn", C.c_long), ("netf", C.c_long),
                ("qlast", C.c_int), ("hlast", C.c_double), ("flag", C.c_int), ("nfails", C.c_int)]


STATS_DTYPE = np.dtype([("nst", "<i8"), ("nfe", "<i8"), ("nfeDQ", "<i8"), ("nsetups", "<i8"),
                        ("nje", "<i8"), ("nni", "<i8"), ("ncfn", "<i8"), ("netf", "<i8"),
                        ("qlast", "<i4"), ("hlast", "<f8"), ("flag", "<i4"), ("nfails", "<i4")], align=True)


def default_phys(**kw) -> Phys:
    """params.dat defaults: zeta 5e-17 s^-1 (normalised by 1.3e-17,
    input_parameters.F90:31), v_turb 1 km/s (3DPDR.F90:47), omega 0.42,
    grain radius 1e-7, metallicity 1, gas-to-dust 100 (params.dat:11-18)."""
    d = dict(zeta=5.0e-17 / 1.3e-17, v_turb=1.0e5, omega=0.42, grain_radius=1.0e-7,
             metallicity=1.0, g2d=100.0)
    d.update(kw)
    return Phys(**d)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build with `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.pdr_net_open.restype = C.c_void_p
        L.pdr_net_open.argtypes = [C.c_char_p]
        L.pdr_net_close.argtypes = [C.c_void_p]
        L.pdr_net_nspec.argtypes = [C.c_void_p]
        L.pdr_net_nreac.argtypes = [C.c_void_p]
        L.pdr_rates.argtypes = [C.c_void_p, C.POINTER(Phys), C.c_double, C.c_double, C.c_int,
                                _dp, _dp, _dp, _dp]
        L.pdr_rates_batch.argtypes = [C.c_void_p, C.POINTER(Phys), C.c_int, _dp, _dp, C.c_int,
                                      _dp, _dp, _dp, _dp]
        L.pdr_rhs.argtypes = [C.c_void_p, _dp, C.c_double, _dp, _dp, C.POINTER(C.c_double)]
        L.pdr_jac.argtypes = [C.c_void_p, _dp, C.c_double, _dp, _dp]
        L.pdr_jac_bordered.argtypes = [C.c_void_p, _dp, C.c_double, _dp, _dp]
        L.pdr_calculate_abundances.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int,
                                               C.c_void_p, C.c_double, C.c_double, C.c_double,
                                               C.c_double, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.pdr_integrate_point.argtypes = [C.c_void_p, _dp, C.c_double, _dp, C.c_double, C.c_int,
                                          _dp, _dp, C.POINTER(Opts), C.POINTER(Stats)]
        L.pdr_test_hybrid_factor.argtypes = [C.c_int, _dp, C.POINTER(C.c_int), C.c_int, C.c_int]
        L.pdr_set_mxstep.argtypes = [C.c_long]
        L.pdr_set_mxstep.restype = None
        L.pdr_set_xe_policy.argtypes = [C.c_int]
        L.pdr_set_xe_policy.restype = None
        L.pdr_integrate_generic.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, _dp,
                                            C.c_double, C.c_int, _dp, _dp, C.POINTER(Opts),
                                            C.POINTER(Stats)]
        _lib = L
    return _lib


class Oracle:
    """One network loaded into the C oracle."""

    def __init__(self, name: str):
        self.name = name
        p = os.path.join(NET_DIR, f"{name}.tab").encode()
        self.h = lib().pdr_net_open(p)
        if not self.h:
            raise RuntimeError(f"cannot load network table {p!r}")
        self.nspec = lib().pdr_net_nspec(self.h)
        self.nreac = lib().pdr_net_nreac(self.h)
        self.neq = self.nspec - 1

    def __del__(self):
        try:
            if self.h:
                lib().pdr_net_close(self.h)
        except Exception:
            pass

    def rates(self, T, Td, G, Av, cols=None, phys: Optional[Phys] = None):
        """Rates of npts points.  G, Av: [npts, nrays]; cols: [npts, nrays, 5]
        (H2, HD, CO, C, S); returns [npts, nreac]."""
        T = np.ascontiguousarray(np.atleast_1d(T), dtype=np.float64)
        npts = T.shape[0]
        Td = np.ascontiguousarray(np.broadcast_to(Td, (npts,)), dtype=np.float64)
        G = np.ascontiguousarray(np.asarray(G, dtype=np.float64).reshape(npts, -1))
        nrays = G.shape[1]
        Av = np.ascontiguousarray(np.asarray(Av, dtype=np.float64).reshape(npts, nrays))
        if cols is None:
            cols = np.zeros((npts, nrays, 5))
        cols = np.ascontiguousarray(np.asarray(cols, dtype=np.float64).reshape(npts, nrays, 5))
        out = np.zeros((npts, self.nreac))
        ph = phys or default_phys()
        nbad = lib().pdr_rates_batch(self.h, C.byref(ph), npts, T, Td, nrays, G.ravel(), Av.ravel(),
                                     cols.ravel(), out.ravel())
        if nbad:
            raise FloatingPointError(f"{nbad} points with a negative rate")
        return out

    def rhs(self, rate, n_H, x):
        rate = np.ascontiguousarray(rate, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        yd = np.zeros(self.neq)
        xe = C.c_double()
        lib().pdr_rhs(self.h, rate, n_H, x, yd, C.byref(xe))
        return yd, xe.value

    def jac(self, rate, n_H, x):
        J = np.zeros(self.neq * self.neq)
        lib().pdr_jac(self.h, np.ascontiguousarray(rate, dtype=np.float64), n_H,
                      np.ascontiguousarray(x, dtype=np.float64), J)
        return J.reshape(self.neq, self.neq)

    def jac_bordered(self, rate, n_H, x):
        """Jacobian of the bordered system, (neq+1)^2: rows < neq hold
        d ydot / d x at fixed x_e and d ydot / d x_e (last column)."""
        nb = self.neq + 1
        J = np.zeros(nb * nb)
        lib().pdr_jac_bordered(self.h, np.ascontiguousarray(rate, dtype=np.float64), n_H,
                               np.ascontiguousarray(x, dtype=np.float64), J)
        return J.reshape(nb, nb)

    def calculate_abundances(self, abundance, rate, density, temperature=None, rtol=1e-8,
                             atol=1e-30, start_time=0.0, end_time=1e7, mode=0,
                             nthreads=0, nelect=None, mxstep=10_000_000, xe_policy=0):
        """calculate_abundances_ semantics; returns (status, abundance, stats).
        mxstep: steps per CVode call; xe_policy 1: electron slot = charge sum
        of the returned abundances (0: x_e of the last RHS evaluation)."""
        ab = np.array(abundance, dtype=np.float64, order="C", copy=True)
        npart = ab.shape[0]
        rate = np.ascontiguousarray(rate, dtype=np.float64)
        density = np.ascontiguousarray(np.broadcast_to(density, (npart,)), dtype=np.float64)
        temp = np.ascontiguousarray(np.broadcast_to(temperature if temperature is not None else 0.0,
                                                    (npart,)), dtype=np.float64)
        stats = np.zeros(npart, dtype=STATS_DTYPE)
        ne = nelect if nelect is not None else self.nspec
        lib().pdr_set_mxstep(int(mxstep))
        lib().pdr_set_xe_policy(int(xe_policy))
        st = lib().pdr_calculate_abundances(ab.ravel(), rate.ravel(), density, temp, npart, self.nspec,
                                            self.nreac, self.h, rtol, atol, start_time, end_time, ne,
                                            mode, nthreads, stats.ctypes.data)
        lib().pdr_set_mxstep(10_000_000)
        lib().pdr_set_xe_policy(0)
        return st, ab, stats

    def integrate(self, rate, n_H, y0, touts, t0=0.0, rtol=1e-8, atol=1e-30, mode=0,
                  mxstep=10_000_000):
        touts = np.ascontiguousarray(touts, dtype=np.float64)
        yout = np.zeros((len(touts), self.neq))
        o = Opts(rtol, atol, mxstep, mode, 5)
        s = Stats()
        flag = lib().pdr_integrate_point(self.h, np.ascontiguousarray(rate, dtype=np.float64), n_H,
                                         np.ascontiguousarray(y0, dtype=np.float64), t0, len(touts),
                                         touts, yout.ravel(), C.byref(o), C.byref(s))
        return flag, yout, s


RHS_FN = C.CFUNCTYPE(C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)
JAC_FN = C.CFUNCTYPE(C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)


def integrate_generic(f, jac, y0, touts, t0=0.0, rtol=1e-8, atol=1e-30, mode=1,
                      mxstep=100000):
    """Run the oracle integrator on a Python-defined ODE (small test problems)."""
    n = len(y0)

    @RHS_FN
    def cf(t, y, yd, ud):
        v = f(t, np.ctypeslib.as_array(y, (n,)).copy())
        for i in range(n):
            yd[i] = v[i]
        return 0

    @JAC_FN
    def cj(t, y, J, ud):
        M = np.asarray(jac(t, np.ctypeslib.as_array(y, (n,)).copy()), dtype=np.float64)
        for i in range(n):
            for j in range(n):
                J[i * n + j] = M[i, j]
        return 0

    touts = np.ascontiguousarray(touts, dtype=np.float64)
    yout = np.zeros((len(touts), n))
    o = Opts(rtol, atol, mxstep, mode, 5)
    s = Stats()
    flag = lib().pdr_integrate_generic(n, C.cast(cf, C.c_void_p), C.cast(cj, C.c_void_p), None,
                                       np.ascontiguousarray(y0, dtype=np.float64), t0, len(touts),
                                       touts, yout.ravel(), C.byref(o), C.byref(s))
    return flag, yout, s


# --------------------------------------------------------------------------
# the reference's own rate code (oracle/_ref, built by `make -C oracle ref`)
# --------------------------------------------------------------------------

def have_ref_rates(name: str) -> bool:
    return os.path.exists(os.path.join(REF_BIN_DIR, f"ref_rates_{name}"))


def ref_rates(name: str, T, Td, G, Av, cols=None, phys: Optional[Phys] = None,
              ref_dir: str = "/root/reference"):
    """Run the amdflang-built reference CALCULATE_REACTION_RATES."""
    T = np.atleast_1d(np.asarray(T, dtype=np.float64))
    npts = T.shape[0]
    Td = np.broadcast_to(np.asarray(Td, dtype=np.float64), (npts,))
    G = np.asarray(G, dtype=np.float64).reshape(npts, -1)
    nrays = G.shape[1]
    Av = np.asarray(Av, dtype=np.float64).reshape(npts, nrays)
    cols = np.zeros((npts, nrays, 5)) if cols is None else np.asarray(cols, dtype=np.float64).reshape(npts, nrays, 5)
    ph = phys or default_phys()
    o = Oracle(name)
    lines = [f"{o.nspec} {o.nreac}",
             " ".join(repr(float(v)) for v in (ph.zeta, ph.v_turb, ph.omega, ph.grain_radius,
                                               ph.metallicity, ph.g2d)),
             f"{npts} {nrays}"]
    for p in range(npts):
        vals = [T[p], Td[p], *G[p], *Av[p], *cols[p].ravel()]
        lines.append(" ".join(repr(float(v)) for v in vals))
    exe = os.path.join(REF_BIN_DIR, f"ref_rates_{name}")
    r = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True,
                       cwd=ref_dir, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"reference rate driver failed: {r.stderr[-2000:]}")
    # READ_RATES also prints warnings to unit 6; keep the numeric lines
    vals = []
    for line in r.stdout.splitlines():
        t = line.strip()
        try:
            vals.append(float(t.replace("D", "E")))
        except ValueError:
            continue
    vals = np.array(vals, dtype=np.float64)
    return vals.reshape(npts, o.nreac)


# --------------------------------------------------------------------------
# ray tracer (raytrace.c) and the reference's own (oracle/_ref/ref_raytrace)
# --------------------------------------------------------------------------

ORDER_HEAPSORT = 0
ORDER_STABLE = 1
AV_FAC = float(np.float32(6.289e-22))   # single-precision literal, input_parameters.F90:57
UV_FAC = float(np.float32(3.02))        # input_parameters.F90:58
PC_CM = 3.08568025e18                   # healpix_types.F90:98


def _raylib():
    L = lib()
    if not getattr(L, "_ray_ready", False):
        L.pdr_hp_pix2vec_nest.argtypes = [C.c_int, C.c_int, _dp]
        L.pdr_hp_vec2pix_nest.argtypes = [C.c_int, _dp]
        L.pdr_eval_points.argtypes = [C.c_int, _dp, C.POINTER(C.c_int), C.c_int, C.c_double, C.c_int, C.c_int,
                                      C.POINTER(C.c_int), C.POINTER(C.c_int), _dp]
        L.pdr_ray_columns.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp, _dp,
                                      C.c_int, _dp, C.c_double, _dp, _dp]
        L.pdr_uvfield.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_double, _dp, _dp, _dp]
        L._ray_ready = True
    return L


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _d(a):
    return a


def hp_vectors(level: int) -> np.ndarray:
    L = _raylib()
    nside = 1 << level
    out = np.zeros((12 * nside * nside, 3))
    for j in range(out.shape[0]):
        v = np.zeros(3)
        L.pdr_hp_pix2vec_nest(nside, j, _d(v))
        out[j] = v
    return out


def hp_vec2pix(level: int, r) -> int:
    r = np.ascontiguousarray(r, dtype=np.float64)
    return int(_raylib().pdr_hp_vec2pix_nest(1 << level, _d(r)))


def eval_points(xyz, level=0, theta_crit=1.3, order=ORDER_STABLE, maxpoints=600, etype=None):
    """Evaluation points of a point cloud (eval_points.F90:129-225).
    Returns (epray[n, nrays], proj[n, nrays, maxpoints+1], epoint[n, nrays, maxpoints+1, 3])."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    nrays = 12 << (2 * level)
    M = maxpoints + 1
    epray = np.zeros((n, nrays), dtype=np.int32)
    proj = np.zeros((n, nrays, M), dtype=np.int32)
    epoint = np.zeros((n, nrays, M, 3))
    et = None if etype is None else np.ascontiguousarray(etype, dtype=np.int32)
    rc = _raylib().pdr_eval_points(n, _d(xyz), None if et is None else _ip(et), level, theta_crit, order,
                                   maxpoints, _ip(epray), _ip(proj), _d(epoint))
    if rc != 0:
        raise RuntimeError("a ray needs more than maxpoints evaluation points")
    return epray, proj, epoint


def ray_columns(epray, proj, epoint, rho, xsel=None, av_fac=AV_FAC):
    """Av[n, nrays] and columns[n, nrays, nsel] along the rays (3DPDR.F90:1650-1664, 1818-1838)."""
    n, nrays, M = proj.shape
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    nsel = 0 if xsel is None else xsel.shape[1]
    xs = np.ascontiguousarray(xsel if xsel is not None else np.zeros((n, 1)), dtype=np.float64)
    Av = np.zeros((n, nrays))
    cols = np.zeros((n, nrays, max(nsel, 1)))
    _raylib().pdr_ray_columns(n, nrays, M - 1, _ip(np.ascontiguousarray(epray)), _ip(np.ascontiguousarray(proj)),
                              _d(np.ascontiguousarray(epoint)), _d(rho), nsel, _d(xs), av_fac, _d(Av), _d(cols))
    return Av, cols[:, :, :nsel]


def uvfield(Av, level=0, gext=(100.0, 0.0, 0.0), fieldchoice=1, uv_fac=UV_FAC):
    """Per-ray surface field G, attenuated field and GUESS_TEMP temperature (3DPDR.F90:1667-1700)."""
    Av = np.ascontiguousarray(Av, dtype=np.float64)
    n, nrays = Av.shape
    G = np.zeros((n, nrays))
    uv = np.zeros(n)
    tg = np.zeros(n)
    g = np.ascontiguousarray(gext, dtype=np.float64)
    _raylib().pdr_uvfield(n, level, fieldchoice, _d(g), _d(Av), uv_fac, _d(G), _d(uv), _d(tg))
    return G, uv, tg


def have_ref_raytrace() -> bool:
    return os.path.exists(os.path.join(REF_BIN_DIR, "ref_raytrace"))


def ref_eval_points(xyz, level=0, theta_crit=1.3, maxpoints=600, etype=None, workdir=None):
    """Run the amdflang-built reference evaluation_points; same return layout as eval_points."""
    import tempfile
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    et = np.ones(n, dtype=int) if etype is None else np.asarray(etype)
    lines = [f"{n} {level} {theta_crit!r} {maxpoints}"]
    lines += [f"{x!r} {y!r} {z!r} {int(t)}" for (x, y, z), t in zip(xyz.tolist(), et.tolist())]
    with tempfile.TemporaryDirectory(dir=workdir) as td:
        r = subprocess.run([os.path.join(REF_BIN_DIR, "ref_raytrace")], input="\n".join(lines) + "\n",
                           capture_output=True, text=True, cwd=td, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"reference ray tracer failed: {r.stdout[-1000:]} {r.stderr[-1000:]}")
        txt = open(os.path.join(td, "raytrace_out.txt")).read().split("\n")
    nrays = 12 << (2 * level)
    M = maxpoints + 1
    vec = np.array([[float(v) for v in txt[j].split()] for j in range(nrays)])
    epray = np.zeros((n, nrays), dtype=np.int32)
    proj = np.zeros((n, nrays, M), dtype=np.int32)
    epoint = np.zeros((n, nrays, M, 3))
    k = nrays
    for p in range(n):
        for j in range(nrays):
            q, jj, cnt = (int(v) for v in txt[k].split())
            k += 1
            epray[p, j] = cnt
            proj[p, j, 0] = p
            epoint[p, j, 0] = xyz[p]
            for i in range(1, cnt + 1):
                f = txt[k].split()
                k += 1
                proj[p, j, i] = int(f[0])
                epoint[p, j, i] = [float(v) for v in f[1:4]]
    return vec, epray, proj, epoint


# --------------------------------------------------------------------------
# heating terms (heating.c) and the reference's CALC_HEATING (oracle/_ref)
# --------------------------------------------------------------------------

HEAT_SPECIES = ("NH", "NH2", "NC", "NCO", "NHEx", "NH3x", "NH3Ox", "NHCOx", "NH2x", "NELECT")
HEAT_REACTIONS = ("NRGR", "NRH2", "NRCI")
HEAT_NET = {"reduced": 0, "full": 1, "mynetwork": 2}


def heating_indices(net) -> np.ndarray:
    """0-based slots the heating reads, from READ_SPECIES's labels (1-based)."""
    return np.array([net.labels[k] - 1 for k in HEAT_SPECIES] + [net.nr_indices[k] - 1 for k in HEAT_REACTIONS],
                    dtype=np.int32)


def heating(name, net, n, T, Td, uv, ab, rate, phys: Optional[Phys] = None) -> np.ndarray:
    L = lib()
    if not getattr(L, "_heat_ready", False):
        L.pdr_heating.argtypes = [C.POINTER(Phys), C.c_int, C.POINTER(C.c_int), C.c_double, C.c_double, C.c_double,
                                  C.c_double, _dp, _dp, _dp]
        L._heat_ready = True
    ph = phys or default_phys()
    idx = heating_indices(net)
    ab = np.ascontiguousarray(ab, dtype=np.float64).reshape(len(n), -1)
    rate = np.ascontiguousarray(rate, dtype=np.float64).reshape(len(n), -1)
    out = np.zeros((len(n), 12))
    for p in range(len(n)):
        h = np.zeros(12)
        L.pdr_heating(C.byref(ph), HEAT_NET[name], idx.ctypes.data_as(C.POINTER(C.c_int)), float(n[p]), float(T[p]),
                      float(Td[p]), float(uv[p]), ab[p], rate[p], h)
        out[p] = h
    return out


def have_ref_heating(name: str) -> bool:
    return os.path.exists(os.path.join(REF_BIN_DIR, f"ref_heating_{name}"))


def ref_heating(name, n, T, Td, uv, ab, rate, phys: Optional[Phys] = None, ref_dir: str = "/root/reference"):
    """Run the amdflang-built reference CALC_HEATING on each point."""
    ph = phys or default_phys()
    o = Oracle(name)
    lines = [f"{o.nspec} {o.nreac}",
             " ".join(repr(float(v)) for v in (ph.zeta, ph.v_turb, ph.omega, ph.grain_radius, ph.metallicity,
                                               ph.g2d)),
             f"{len(n)}"]
    for p in range(len(n)):
        lines.append(" ".join(repr(float(v)) for v in (n[p], T[p], Td[p], uv[p])))
        lines.append(" ".join(repr(float(v)) for v in ab[p]))
        lines.append(" ".join(repr(float(v)) for v in rate[p]))
    r = subprocess.run([os.path.join(REF_BIN_DIR, f"ref_heating_{name}")], input="\n".join(lines) + "\n",
                       capture_output=True, text=True, cwd=ref_dir, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"reference heating driver failed: {r.stderr[-2000:]}")
    vals = []
    for line in r.stdout.splitlines():
        try:
            vals.append(float(line.strip().replace("D", "E")))
        except ValueError:
            continue
    return np.array(vals).reshape(len(n), 12)


# --------------------------------------------------------------------------
# level populations, line cooling, dust temperature (thermal.c) and the
# reference's own routines (oracle/_ref/ref_lvg_{1d,3d})
# --------------------------------------------------------------------------

class Coolant(C.Structure):
    _fields_ = [("nlev", C.c_int), ("ntemp", C.c_int), ("E", C.c_void_p), ("g", C.c_void_p), ("A", C.c_void_p),
                ("B", C.c_void_p), ("F", C.c_void_p), ("temps", C.c_void_p), ("col", C.c_void_p)]


def _thermlib():
    L = lib()
    if not getattr(L, "_therm_ready", False):
        L.pdr_find_ccoeff.argtypes = [C.POINTER(Coolant), C.c_double, _dp, _dp]
        L.pdr_escape_probability.argtypes = [C.POINTER(Coolant), C.c_double, C.c_double, C.c_double, C.c_double,
                                             C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int),
                                             C.POINTER(C.c_longlong), _dp, C.POINTER(C.c_longlong), _dp, _dp, _dp,
                                             _dp, _dp, _dp, _dp]
        L.pdr_escape_probability.restype = C.c_double
        L.pdr_solvlevpop.argtypes = [C.c_int, _dp, C.c_double, _dp]
        L.pdr_lte_populations.argtypes = [C.POINTER(Coolant), C.c_double, C.c_double, _dp]
        L.pdr_dust_temperature.argtypes = [C.c_double, C.c_int, _dp]
        L.pdr_dust_temperature.restype = C.c_double
        L._therm_ready = True
    return L


class CoolantTables:
    """a pdrchem.lamda.Coolant held in the C layout the oracle reads"""

    def __init__(self, cool):
        self.cool = cool
        self.arrays = cool.fortran_arrays()
        E, g, A, B, F, TT, col = self.arrays
        self.s = Coolant(cool.nlev, cool.ntemp, *[a.ctypes.data for a in (E, g, A, B, F, TT, col)])


def find_ccoeff(tab: CoolantTables, T, dens):
    n = tab.cool.nlev
    out = np.zeros(n * n)
    _thermlib().pdr_find_ccoeff(C.byref(tab.s), float(T), np.ascontiguousarray(dens, dtype=np.float64), out)
    return out.reshape(n, n, order="F")


def lvg_point(tab: CoolantTables, T, Td, rho, metallicity, v_turb, g2d, dens, xspec, pop, chains, pseudo_ray=-1):
    """Level populations of one point and coolant.  chains: per ray a tuple
    (pops [m+1, nlev] with row 0 the point's own, steps [m] in pc).  Returns a
    dict with C, transition, cooling, line, tau, beta (nlev x nlev [x nrays],
    pairs i > j filled) and solution."""
    L = _thermlib()
    n = tab.cool.nlev
    nrays = len(chains)
    epray = np.array([len(s) for _, s in chains], dtype=np.int32)
    pop_off = np.zeros(nrays, dtype=np.int64)
    step_off = np.zeros(nrays, dtype=np.int64)
    for j in range(1, nrays):
        pop_off[j] = pop_off[j - 1] + epray[j - 1] + 1
        step_off[j] = step_off[j - 1] + epray[j - 1]
    cp = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.float64).reshape(-1, n) for p, _ in chains]))
    st = np.ascontiguousarray(np.concatenate([np.asarray(s, dtype=np.float64) for _, s in chains] + [np.zeros(1)]))
    Cc = find_ccoeff(tab, T, dens)
    Cf = np.ascontiguousarray(Cc.ravel(order="F"))
    tr = np.zeros(n * n)
    line = np.zeros(n * n)
    tau = np.zeros(n * n * nrays)
    beta = np.zeros(n * n * nrays)
    popv = np.ascontiguousarray(pop, dtype=np.float64)
    cool = L.pdr_escape_probability(C.byref(tab.s), float(Td), float(T), float(v_turb), float(rho), float(metallicity),
                                    float(g2d), nrays, int(pseudo_ray), epray.ctypes.data_as(C.POINTER(C.c_int)),
                                    pop_off.ctypes.data_as(C.POINTER(C.c_longlong)), cp,
                                    step_off.ctypes.data_as(C.POINTER(C.c_longlong)), st, popv, Cf, tr, line, tau,
                                    beta)
    sol = np.zeros(n)
    L.pdr_solvlevpop(n, tr, float(xspec * rho), sol)
    F = lambda a, *sh: a.reshape(*sh, order="F")
    return {"C": Cc, "transition": F(tr, n, n), "cooling": cool, "line": F(line, n, n),
            "tau": F(tau, n, n, nrays), "beta": F(beta, n, n, nrays), "solution": sol}


def lte_populations(tab: CoolantTables, T, density):
    out = np.zeros(tab.cool.nlev)
    _thermlib().pdr_lte_populations(C.byref(tab.s), float(T), float(density), out)
    return out


def dust_temperature(uv, rad_surface):
    rs = np.ascontiguousarray(rad_surface, dtype=np.float64)
    return float(_thermlib().pdr_dust_temperature(float(uv), len(rs), rs))


def have_ref_lvg() -> bool:
    return all(os.path.exists(os.path.join(REF_BIN_DIR, f"ref_lvg_{k}")) for k in ("1d", "3d"))


def _ref_lvg_run(variant, text):
    r = subprocess.run([os.path.join(REF_BIN_DIR, f"ref_lvg_{variant}")], input=text, capture_output=True, text=True,
                       timeout=600, cwd="/tmp")
    if r.returncode != 0:
        raise RuntimeError(f"reference level-population driver failed: {r.stdout[-1000:]} {r.stderr[-1000:]}")
    vals = []
    for line in r.stdout.splitlines():
        try:
            vals.append(float(line.strip().replace("D", "E")))
        except ValueError:
            continue
    return np.array(vals)


def ref_coolant_tables(path, nlev, ntemp):
    """READINPUT of the reference on a coolant file: (E, g, A, B, F, TT,
    col[7]) with col in the partner order H2, pH2, oH2, e, H, He, H+"""
    v = _ref_lvg_run("3d", f"1\n{path}\n{nlev} {ntemp}\n")
    k = 0

    def take(m, *sh):
        nonlocal k
        a = v[k:k + m].reshape(*sh, order="F") if sh else v[k:k + m]
        k += m
        return a
    n2, n3 = nlev * nlev, nlev * nlev * ntemp
    E, g = take(nlev), take(nlev)
    A, B, F = take(n2, nlev, nlev), take(n2, nlev, nlev), take(n2, nlev, nlev)
    TT = take(7 * ntemp, 7, ntemp)
    col = np.stack([take(n3, nlev, nlev, ntemp) for _ in range(7)])
    return E, g, A, B, F, TT, col


def ref_lvg(path, nlev, ntemp, coolant_id, cases, pseudo_1d):
    """The reference's FIND_CCOEFF, escape_probability and solvlevpop on
    cases: dicts with T, Td, rho, metallicity, v_turb, g2d, dens (5 partner
    abundances, fractional), xspec, pop, and rays: per ray (epoints [m+1, 3],
    pops [m+1, nlev]) with row 0 the point itself.  Returns per case the dict
    lvg_point returns."""
    lines = ["2", str(path), f"{nlev} {ntemp} {coolant_id} {len(cases)}"]
    for c in cases:
        nrays = len(c["rays"])
        lines.append(" ".join(repr(float(x)) for x in (c["T"], c["Td"], c["rho"], c["metallicity"], c["v_turb"],
                                                        c["g2d"])) + f" {nrays}")
        lines.append(" ".join(repr(float(x)) for x in list(c["dens"]) + [c["xspec"]]))
        lines.append(" ".join(repr(float(x)) for x in c["pop"]))
        for ept, pops in c["rays"]:
            lines.append(str(len(ept) - 1))
            for i in range(len(ept)):
                lines.append(" ".join(repr(float(x)) for x in list(ept[i]) + list(pops[i])))
    v = _ref_lvg_run("1d" if pseudo_1d else "3d", "\n".join(lines) + "\n")
    out, k, n2 = [], 0, nlev * nlev
    for c in cases:
        nrays = len(c["rays"])
        d = {}
        d["C"] = v[k:k + n2].reshape(nlev, nlev, order="F"); k += n2
        d["transition"] = v[k:k + n2].reshape(nlev, nlev, order="F"); k += n2
        d["cooling"] = v[k]; k += 1
        d["line"] = v[k:k + n2].reshape(nlev, nlev, order="F"); k += n2
        d["tau"] = v[k:k + n2 * nrays].reshape(nlev, nlev, nrays, order="F"); k += n2 * nrays
        d["beta"] = v[k:k + n2 * nrays].reshape(nlev, nlev, nrays, order="F"); k += n2 * nrays
        d["solution"] = v[k:k + nlev]; k += nlev
        out.append(d)
    return out


def ref_dust_temperatures(uv, rad_surface, av_fac=AV_FAC):
    rad_surface = np.asarray(rad_surface, dtype=np.float64)
    n, nrays = rad_surface.shape
    lines = ["3", f"{n} {nrays} {av_fac!r}"]
    for p in range(n):
        lines.append(" ".join(repr(float(x)) for x in [uv[p]] + list(rad_surface[p])))
    return _ref_lvg_run("3d", "\n".join(lines) + "\n")
