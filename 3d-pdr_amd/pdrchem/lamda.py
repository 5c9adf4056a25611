"""Coolant data files in the LAMDA/RADEX format (12c+.dat, 12c.dat, 16o.dat,
12co.dat), read with the semantics of the reference's READINPUT
(read_input.F90:6-259), into the arrays the thermal balance uses.

Arrays are kept in the reference's (Fortran, column-major) layout so that a
Fortran host can hand its own module arrays to the library unchanged:
A(i, j) of a nlev x nlev matrix is element i + j*nlev (0-based i, j);
COL(i, j, k) is i + j*nlev + k*nlev*nlev; TEMPERATURES(p, k) is p + 7*k.

Partners (the file's partner id): 1 H2, 2 para-H2, 3 ortho-H2, 4 electrons,
5 H, 6 He, 7 H+.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# cgs constants of healpix_types.F90:86-89
KB = 1.38065040e-16
C_LIGHT = 2.99792458e10
HP = 6.62606896e-27

NPARTNERS = 7

# the reference's sizes (input_parameters.F90:120-127) and its files
# (params.dat): coolant order CII, CI, OI, CO
COOLANTS = (("CII", "12c+.dat", 5, 18), ("CI", "12c.dat", 5, 29), ("OI", "16o.dat", 5, 27),
            ("CO", "12co.dat", 41, 25))


@dataclass
class Coolant:
    name: str
    nlev: int
    ntemp: int
    energies: np.ndarray      # [nlev] erg
    weights: np.ndarray       # [nlev]
    A: np.ndarray             # [nlev, nlev] s^-1, A[i, j] for upper i -> lower j
    B: np.ndarray             # [nlev, nlev]
    freq: np.ndarray          # [nlev, nlev] Hz
    temperatures: np.ndarray  # [7, ntemp] K, zero where the partner has no data
    col: np.ndarray           # [7, nlev, nlev, ntemp] cm^3 s^-1, partner p at index p-1

    def fortran_arrays(self):
        """(energies, weights, A, B, freq, temperatures, col) flattened in the
        reference's column-major order; col as 7 blocks of nlev*nlev*ntemp"""
        f = lambda a: np.ascontiguousarray(a.ravel(order="F"))
        cols = np.concatenate([f(self.col[p]) for p in range(NPARTNERS)])
        return (np.ascontiguousarray(self.energies), np.ascontiguousarray(self.weights), f(self.A), f(self.B),
                f(self.freq), f(self.temperatures), cols)


class _Records:
    """list-directed READ(unit,*): each READ starts at a new line and takes as
    many values as it needs from that line and the following ones; what is
    left of its last line is discarded"""

    def __init__(self, lines):
        self.lines = lines
        self.pos = 0

    def skip(self, n=1):
        self.pos += n

    def read(self, n):
        vals = []
        while len(vals) < n:
            if self.pos >= len(self.lines):
                raise ValueError("end of file inside a READ")
            toks = self.lines[self.pos].replace(",", " ").split()
            self.pos += 1
            for t in toks:
                if t.startswith("/"):
                    break
                vals.append(t)
                if len(vals) == n:
                    break
        return vals


def _num(t):
    return float(t.replace("D", "E").replace("d", "e"))


def read_coolant(path: str, nlev: int, ntemp: int, name: str = "") -> Coolant:
    """READINPUT(FILENAME, NLEV, NTEMP, ...) (read_input.F90:6-259)."""
    with open(path) as fh:
        R = _Records(fh.read().splitlines())
    E = np.zeros(nlev)
    g = np.zeros(nlev)
    A = np.zeros((nlev, nlev))
    B = np.zeros((nlev, nlev))
    F = np.zeros((nlev, nlev))
    TT = np.zeros((NPARTNERS, ntemp))
    COL = np.zeros((NPARTNERS, nlev, nlev, ntemp))
    R.skip(5)                                    # READ(8,'(////)')
    n = int(R.read(1)[0])
    if n != nlev:
        raise ValueError(f"{path}: {n} energy levels, expected {nlev}")
    R.skip()
    n = 0
    while n < nlev:
        v = R.read(3)
        n = int(v[0])
        E[n - 1] = _num(v[1]) * C_LIGHT * HP     # cm^-1 -> erg
        g[n - 1] = _num(v[2])
    R.skip()
    nlin = int(R.read(1)[0])
    R.skip()
    n = 0
    while n < nlin:
        v = R.read(5)
        n, i, j = int(v[0]), int(v[1]) - 1, int(v[2]) - 1
        F[i, j] = _num(v[4]) * 1.0e9             # GHz -> Hz
        F[j, i] = F[i, j]
        A[i, j] = _num(v[3])
        B[i, j] = A[i, j] / (2.0 * HP * (F[i, j] ** 3) / (C_LIGHT ** 2))
        B[j, i] = B[i, j] * (g[i] / g[j])
    for i in range(nlev):
        for j in range(nlev):
            fr = abs(E[i] - E[j]) / HP
            if F[i, j] != 0.0:
                if abs(fr - F[i, j]) / F[i, j] > 1.0e-2:
                    raise ValueError(f"{path}: calculated frequency differs by >1% ({fr} vs {F[i, j]} Hz)")
            else:
                F[i, j] = fr
    R.skip()
    npart = int(R.read(1)[0])
    for _ in range(npart):
        R.skip()
        p = int(R.read(1)[0])
        R.skip()
        ncol = int(R.read(1)[0])
        R.skip()
        m = int(R.read(1)[0])
        if m > ntemp:
            raise ValueError(f"{path}: {m} temperatures, more than {ntemp}")
        if not 1 <= p <= NPARTNERS:
            raise ValueError(f"{path}: unknown collision partner {p}")
        R.skip()
        TT[p - 1, :m] = [_num(t) for t in R.read(m)]
        R.skip()
        n = 0
        while n < ncol:
            v = R.read(3 + m)
            n, i, j = int(v[0]), int(v[1]) - 1, int(v[2]) - 1
            for k in range(m):
                COL[p - 1, i, j, k] = _num(v[3 + k])
                # reverse (excitation) rate by detailed balance where the
                # file has given none (so far)
                if COL[p - 1, i, j, k] != 0.0 and COL[p - 1, j, i, k] == 0.0:
                    COL[p - 1, j, i, k] = COL[p - 1, i, j, k] * (g[i] / g[j]) * \
                        np.exp(-(E[i] - E[j]) / (KB * TT[p - 1, k]))
    return Coolant(name, nlev, ntemp, E, g, A, B, F, TT, COL)


def read_coolants(directory: str, files=None):
    """The four coolants of the reference (CII, CI, OI, CO), from `directory`."""
    import os
    out = []
    for name, fn, nlev, ntemp in COOLANTS:
        f = (files or {}).get(name, fn)
        out.append(read_coolant(os.path.join(directory, f), nlev, ntemp, name))
    return out


COOLANT_DIR = __import__("os").path.join(__import__("os").path.dirname(__import__("os").path.dirname(
    __import__("os").path.abspath(__file__))), "coolants")


def generate_all(ref_dir: str, out_dir: str = COOLANT_DIR) -> None:
    """Parse the reference's four coolant files and store the tables (the
    parsed data, as READINPUT leaves them) as <name>.npz in out_dir."""
    import os
    os.makedirs(out_dir, exist_ok=True)
    for c in read_coolants(ref_dir):
        np.savez_compressed(os.path.join(out_dir, f"{c.name}.npz"), nlev=c.nlev, ntemp=c.ntemp, energies=c.energies,
                            weights=c.weights, A=c.A, B=c.B, freq=c.freq, temperatures=c.temperatures, col=c.col)


def load_coolant(name: str, directory: str = COOLANT_DIR) -> Coolant:
    import os
    d = np.load(os.path.join(directory, f"{name}.npz"))
    return Coolant(name, int(d["nlev"]), int(d["ntemp"]), d["energies"], d["weights"], d["A"], d["B"], d["freq"],
                   d["temperatures"], d["col"])


def load_coolants(directory: str = COOLANT_DIR):
    """the four coolants (CII, CI, OI, CO) from the stored tables"""
    return [load_coolant(name, directory) for name, _, _, _ in COOLANTS]
