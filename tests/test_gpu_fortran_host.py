"""The drop-in boundary from Fortran: a host program built with amdflang on
the reference's own modules.F90 (so the tolerances, start/end time and
electron index are the Fortran modules' bind(c) variables, modules.F90:165-166,
383, 529-530) calls CALCULATE_ABUNDANCES with the reference's implicit
interface (3DPDR.F90:498) and links libpdrchem.so in place of
calculate_abundances.o and SUNDIALS (oracle/ref_host_driver.F90, built by
`make -C oracle ref`).  Its results equal the same call made through ctypes,
follow the end time set in the Fortran module, and meet the parity bar
against the FP64 oracle."""
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
from pdrchem import lib, workloads
from pdrchem.netgen import load_network

pytestmark = pytest.mark.gpu
HOST = os.path.join(os.path.dirname(oracle.__file__), "_ref", "fortran_host")


def run_host(x0, rate, nH, T, nelect, t_end, rtol=1e-8, atol=1e-30):
    npart, nspec = x0.shape
    nreac = rate.shape[1]
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("4i", npart, nspec, nreac, nelect))
            f.write(struct.pack("4d", rtol, atol, 0.0, t_end))
            for a in (x0, rate, nH, T):
                f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        r = subprocess.run([HOST, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr[-2000:]
        b = open(fout, "rb").read()
    status = struct.unpack("i", b[:4])[0]
    return status, np.frombuffer(b[4:], dtype=np.float64).reshape(npart, nspec).copy()


@pytest.mark.skipif(not os.path.exists(HOST), reason="Fortran host not built (make -C oracle ref)")
def test_fortran_host_calls_the_library():
    print("fortran host: 16 slab points, reduced network", flush=True)
    net = load_network("reduced")
    P = workloads.slab_1d(1000).subset(np.arange(0, 1000, 64))
    n = P.n
    x0 = workloads.initial_abundances(net, n)
    o = oracle.Oracle("reduced")
    rate = o.rates(P.T, P.Td, P.G, P.Av, P.cols)
    nelect = net.nspec                       # electrons last (1-based index)
    st, got = run_host(x0, rate, P.n_H, P.T, nelect, 1e7)
    assert st == 0
    # the same call through ctypes, globals set from Python
    lib.set_globals(rtol=1e-8, atol=1e-30, start_time=0.0, end_time=1e7, nelect=nelect)
    ab = x0.copy()
    assert lib.calculate_abundances(ab, np.ascontiguousarray(rate), P.n_H.copy(), P.T.copy()) == 0
    assert np.array_equal(got, ab)
    # parity against the FP64 oracle
    s_, ref, _ = o.calculate_abundances(x0.copy(), rate, P.n_H, P.T, mode=oracle.MODE_GPU)
    m = ref[:, :-1] > 1e-20
    err = float(np.max(np.abs(got[:, :-1][m] - ref[:, :-1][m]) / ref[:, :-1][m]))
    print(f"  max rel err vs oracle {err:.2e}", flush=True)
    assert s_ == 0 and err <= 1e-4
    # the end time is the one in the Fortran module
    st2, got2 = run_host(x0, rate, P.n_H, P.T, nelect, 1e5)
    lib.set_globals(end_time=1e5)
    ab2 = x0.copy()
    lib.calculate_abundances(ab2, np.ascontiguousarray(rate), P.n_H.copy(), P.T.copy())
    lib.set_globals(end_time=1e7)
    assert st2 == 0 and np.array_equal(got2, ab2) and not np.array_equal(got2, got)
