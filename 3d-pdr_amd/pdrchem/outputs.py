"""Output files in the reference's formats.

``write_pdr_fin`` writes the abundance file ``<output>.pdr.fin`` as
3DPDR.F90:1270-1305 does: one line per PDR point, Fortran edit descriptors
'(I7,5ES11.3,I5,300ES11.3)' for 3D runs -- point id, x, y, z, gas and dust
temperature, element type, n_H, attenuated UV field, the abundances, Av per
ray -- and '(I7,4ES11.3,I5,300ES11.3)' with -DPSEUDO_1D -- id, x, Av of ray 6,
gas and dust temperature, element type, n_H, UV field, abundances.  The
reference swaps its last two PDR points on input and back on output
(3DPDR.F90:206-232, 1284-1305); points here are written in input order,
which is the same file.  The cooling, heating, line and level-population
files need the thermal balance, which this library does not run.
"""
from __future__ import annotations

import numpy as np


def _es(v: float, w: int = 11, d: int = 3) -> str:
    """Fortran ESw.d: scientific with one leading digit; exponents beyond two
    digits drop the 'E', as Fortran prints them."""
    if v == 0.0 or not np.isfinite(v):
        s = f"{0.0:.{d}E}" if v == 0.0 else f"{v}"
    else:
        s = f"{v:.{d}E}"
        mant, exp = s.split("E")
        e = int(exp)
        if abs(e) > 99:
            s = f"{mant}{e:+04d}"
    return s.rjust(w)


def pdr_fin_line(pid: int, xyz, tgas: float, tdust: float, etype: int, rho: float, uv: float, abundance, av,
                 pseudo_1d: bool = False) -> str:
    if pseudo_1d:
        head = [xyz[0], av[6], tgas, tdust]
        tail = [rho, uv, *abundance]
    else:
        head = [xyz[0], xyz[1], xyz[2], tgas, tdust]
        tail = [rho, uv, *abundance, *av]
    return f"{pid:7d}" + "".join(_es(float(v)) for v in head) + f"{int(etype):5d}" + "".join(
        _es(float(v)) for v in tail)


def write_pdr_fin(path: str, xyz, tgas, tdust, rho, uv, abundance, av, etype=None, pseudo_1d: bool = False) -> None:
    """xyz [n, 3] pc; tgas, tdust, rho, uv [n]; abundance [n, nspec]; av [n, nrays]."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    et = np.ones(n, dtype=int) if etype is None else np.asarray(etype)
    with open(path, "w") as f:
        for p in range(n):
            f.write(pdr_fin_line(p + 1, xyz[p], tgas[p], tdust[p], et[p], rho[p], uv[p], abundance[p], av[p],
                                 pseudo_1d) + "\n")


def write_cloud_pdr_fin(path: str, cloud, xyz, pseudo_1d: bool = False) -> None:
    """A grid or point cloud's current state (pdrchem.cloud.Cloud) as <output>.pdr.fin."""
    from . import cloud as CL
    n = cloud.npts
    T = cloud.get(CL.TEMPERATURE)
    write_pdr_fin(path, xyz, T, np.full(n, cloud.dust_temperature), cloud.density, cloud.get(CL.UVFIELD),
                  cloud.get(CL.ABUNDANCE), cloud.get(CL.AV), pseudo_1d=pseudo_1d)
