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
                ("qlast", C.c_int), ("hlast", C.c_double), ("flag", C.c_int)]


STATS_DTYPE = np.dtype([("nst", "<i8"), ("nfe", "<i8"), ("nfeDQ", "<i8"), ("nsetups", "<i8"),
                        ("nje", "<i8"), ("nni", "<i8"), ("ncfn", "<i8"), ("netf", "<i8"),
                        ("qlast", "<i4"), ("hlast", "<f8"), ("flag", "<i4")], align=True)


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
                             nthreads=0, nelect=None):
        """calculate_abundances_ semantics; returns (status, abundance, stats)."""
        ab = np.array(abundance, dtype=np.float64, order="C", copy=True)
        npart = ab.shape[0]
        rate = np.ascontiguousarray(rate, dtype=np.float64)
        density = np.ascontiguousarray(np.broadcast_to(density, (npart,)), dtype=np.float64)
        temp = np.ascontiguousarray(np.broadcast_to(temperature if temperature is not None else 0.0,
                                                    (npart,)), dtype=np.float64)
        stats = np.zeros(npart, dtype=STATS_DTYPE)
        ne = nelect if nelect is not None else self.nspec
        st = lib().pdr_calculate_abundances(ab.ravel(), rate.ravel(), density, temp, npart, self.nspec,
                                            self.nreac, self.h, rtol, atol, start_time, end_time, ne,
                                            mode, nthreads, stats.ctypes.data)
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
