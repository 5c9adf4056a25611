"""Synthetic grid-point workloads (physical inputs of the chemistry solve).

slab_1d        1-D slab of 1,000 points: n_H = 1e3 cm^-3, T_gas linear
               10-300 K, A_V linear 0-10 mag, one ray with G0 = 10 Draine,
               zeta = 5e-17 s^-1, T_dust = 20 K (params.dat:14), zero column
               densities.
random_points  seeded points: log10 n_H ~ U[2,6], log10 T ~ U[1,3],
               A_V ~ U[0,20], log10 G0 ~ U[0,5]; optionally log10 N(H2),
               log10 N(CO) ~ U[14,22] cm^-2.
initial_abundances  H = 1.0 (atomic), He = 0.1, C+ = 1.4e-4, O = 3.0e-4, the
               rest 1e-30; e- = sum of ion abundances (read_species.F90:178-190).

The random stream is a counter-based hash (splitmix64 of seed, stream and
point index), so any subset of points can be regenerated independently and
CPU and GPU consumers see identical inputs.
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np

from .netgen import Network

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, index: np.ndarray) -> np.ndarray:
    """U[0,1) doubles for (seed, stream, index); 53-bit resolution."""
    with np.errstate(over="ignore"):
        k = (np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream) * np.uint64(0x9E3779B1)) & _M64
        h = _splitmix64(_splitmix64(index.astype(np.uint64) ^ k))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclasses.dataclass
class Points:
    n_H: np.ndarray       # [n]
    T: np.ndarray         # [n]
    Td: np.ndarray        # [n]
    G: np.ndarray         # [n, nrays]
    Av: np.ndarray        # [n, nrays]
    cols: np.ndarray      # [n, nrays, 5]  H2, HD, CO, C, S

    @property
    def n(self) -> int:
        return self.n_H.shape[0]

    def subset(self, idx) -> "Points":
        return Points(self.n_H[idx], self.T[idx], self.Td[idx], self.G[idx], self.Av[idx], self.cols[idx])


def slab_1d(n: int = 1000) -> Points:
    T = np.linspace(10.0, 300.0, n)
    Av = np.linspace(0.0, 10.0, n)
    return Points(n_H=np.full(n, 1.0e3), T=T, Td=np.full(n, 20.0), G=np.full((n, 1), 10.0),
                  Av=Av[:, None].copy(), cols=np.zeros((n, 1, 5)))


def random_points(n: int, seed: int = 20251203, start: int = 0, columns: bool = False) -> Points:
    idx = np.arange(start, start + n, dtype=np.uint64)
    n_H = 10.0 ** (2.0 + 4.0 * uniform(seed, 1, idx))
    T = 10.0 ** (1.0 + 2.0 * uniform(seed, 2, idx))
    Av = 20.0 * uniform(seed, 3, idx)
    G0 = 10.0 ** (5.0 * uniform(seed, 4, idx))
    cols = np.zeros((n, 1, 5))
    if columns:
        cols[:, 0, 0] = 10.0 ** (14.0 + 8.0 * uniform(seed, 5, idx))  # N(H2)
        cols[:, 0, 2] = 10.0 ** (14.0 + 8.0 * uniform(seed, 6, idx))  # N(CO)
    return Points(n_H=n_H, T=T, Td=np.full(n, 20.0), G=G0[:, None].copy(), Av=Av[:, None].copy(), cols=cols)


def initial_abundances(net: Network, n: int, rest: float = 1.0e-30) -> np.ndarray:
    names = [s.name for s in net.species]
    x = np.full(net.nspec, rest)
    for nm, v in (("H", 1.0), ("He", 0.1), ("C+", 1.4e-4), ("O", 3.0e-4)):
        x[names.index(nm)] = v
    charges = np.array([s.charge for s in net.species], dtype=np.float64)
    ion = charges[:-1] > 0
    x[-1] = float(np.sum(x[:-1][ion] * charges[:-1][ion]))  # e- from the charge sum
    return np.tile(x, (n, 1))


def zeta_normalised(zeta_s: float = 5.0e-17) -> float:
    """Cosmic-ray rate in the units the rate code uses (input_parameters.F90:31)."""
    return zeta_s / 1.3e-17


@dataclasses.dataclass
class Request:
    arrival: float        # seconds after the start of the stream
    network: str
    npts: int
    t_end: float          # years
    first: int            # index of the request's first point in its network's point stream


def request_stream(duration: float = 60.0, rate: float = 200.0, seed: int = 20251204,
                   networks=("reduced", "full")) -> list:
    """Serving workload: Poisson arrivals at `rate` per second for `duration`
    seconds; points per request floor(exp(U[0, ln 4096])) (1..4096, mean
    ~492); network chosen with equal probability; end time log-uniform in
    [1e3, 1e7] yr.  Each network's requests take consecutive points of that
    network's seeded random-point stream (random_points(..., start=first))."""
    kmax = int(duration * rate * 1.5) + 64
    idx = np.arange(kmax, dtype=np.uint64)
    gaps = -np.log(1.0 - uniform(seed, 10, idx)) / rate
    arrival = np.cumsum(gaps)
    which = (uniform(seed, 11, idx) * len(networks)).astype(np.int64)
    npts = np.clip(np.floor(np.exp(uniform(seed, 12, idx) * np.log(4096.0))).astype(np.int64), 1, 4096)
    t_end = 10.0 ** (3.0 + 4.0 * uniform(seed, 13, idx))
    out = []
    firsts = {nm: 0 for nm in networks}
    for k in range(kmax):
        if arrival[k] >= duration:
            break
        nm = networks[which[k]]
        out.append(Request(float(arrival[k]), nm, int(npts[k]), float(t_end[k]), firsts[nm]))
        firsts[nm] += int(npts[k])
    return out


def cloud_density(side: int = 64, mean: float = 1.0e3, sigma_ln: float = 1.0, corr_cells: float = 4.0,
                  seed: int = 20251205) -> np.ndarray:
    """Lognormal n_H on a side^3 grid (index ix + side*(iy + side*iz)):
    ln n_H is a Gaussian random field (white noise from the counter-based
    generator, smoothed by a periodic Gaussian of ``corr_cells`` cells and
    rescaled to standard deviation ``sigma_ln``) with mean ln(mean) -
    sigma_ln^2/2, so that the mean of n_H is ``mean``."""
    n = side ** 3
    idx = np.arange(n, dtype=np.uint64)
    u1 = np.clip(uniform(seed, 11, idx), 1e-300, 1.0)
    u2 = uniform(seed, 12, idx)
    w = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    w = w.reshape(side, side, side)           # [iz, iy, ix]
    k = np.fft.fftfreq(side)
    kk = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    f = np.real(np.fft.ifftn(np.fft.fftn(w) * np.exp(-2.0 * (np.pi * corr_cells) ** 2 * kk)))
    f = (f - f.mean()) / f.std()
    return np.exp(np.log(mean) - 0.5 * sigma_ln ** 2 + sigma_ln * f).reshape(-1)
