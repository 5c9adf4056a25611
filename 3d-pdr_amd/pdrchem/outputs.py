"""Output files in the reference's formats (3DPDR.F90:1270-1500).

One line per PDR point, Fortran edit descriptors as the reference writes them:

* ``<output>.pdr.fin`` (3DPDR.F90:1270-1303): '(I7,5ES11.3,I5,300ES11.3)' in
  3D -- point id, x, y, z, gas temperature (previousgastemperature), dust
  temperature, element type, n_H, attenuated UV field, the abundances, Av per
  ray -- and '(I7,4ES11.3,I5,300ES11.3)' with PSEUDO_1D -- id, x, Av of ray 6,
  gas and dust temperature, element type, n_H, UV field, abundances;
* ``.cool.fin`` (:1343-1358) '(I7,200ES11.3)': id, x[, y, z | Av(6)], CII, CI,
  OI, CO cooling and their sum[, Av per ray in 3D];
* ``.heat.fin`` (:1366-1381) '(I7,200ES11.3)': id, x[, y, z | Av(6)], the 12
  heating terms[, Av per ray];
* ``.line.fin`` (:1389-1449) '(I9,200ES11.3)': index, x[, y, z | Av(6)], the
  17 line emissivities (CII 158um; CI 2-1, 3-1, 3-2; OI 2-1, 3-1, 3-2; CO
  J=1-0 .. 10-9)[, Av per ray];
* ``.spop.fin`` (:1457-1468) '(I9,200ES11.3)': index, Av(6), the level
  populations of CII, CI, OI, CO;
* ``.opdp.fin`` (:1484-1500, PSEUDO_1D only) '(I9,200ES11.3)': index, x, the
  17 lines' optical depths along ray 6 (the reference's 2017 benchmark files
  hold Av(6) in the second column; its source writes x, as here).

Row order.  The reference swaps the data of its last two PDR points when it
reads them (3DPDR.F90:206-232) and computes in that "internal" order.  The
.pdr.fin and .line.fin files write the internal points 1..n-2, n, n-1 --
their rows come out in input order -- while .cool, .heat, .spop and .opdp
write 1..n, so their last two rows hold the last two input points swapped.
The id column is the point's input line number (IDlist_pdr) in .pdr, .cool
and .heat, and the internal index pp in .line, .spop and .opdp.  The writers
here take arrays in the internal order (a cloud created from the points in
``internal_order``) and reproduce both conventions.

Abundance columns.  The reference's arrays are filled by its generated ODE
code, whose reduced network numbers slots 3-5 as CH5+, O2, H2+ where
species_reduced.d lists H2+, CH5+, O2 (odes_reduced.c: the loss terms of
ode[4] are O2's reactions, the electron sum counts x[5]); its .pdr.fin
therefore holds CH5+, O2, H2+ in the columns the species file labels H2+,
CH5+, O2 (Benchmarks B01-B04 show it).  This library keeps species-file
order; ``write_model_outputs`` writes the reference's column order
(``reference_columns``) so that its files compare column for column.
"""
from __future__ import annotations

import math

import numpy as np

NLINES = 17
AV_RAY = 6  # the ray whose Av the PSEUDO_1D files print (3DPDR.F90:1277-1500)


def _es(v: float, w: int = 11, d: int = 3) -> str:
    """Fortran ESw.d: scientific with one leading digit; exponents beyond two
    digits drop the 'E', as Fortran prints them; NaN and infinities as
    gfortran writes them."""
    v = float(v)
    if math.isnan(v):
        s = "NaN"
    elif math.isinf(v):
        s = "Infinity" if v > 0 else "-Infinity"
    elif v == 0.0:
        s = f"{0.0:.{d}E}"
    else:
        s = f"{v:.{d}E}"
        mant, exp = s.split("E")
        e = int(exp)
        if abs(e) > 99:
            s = f"{mant}{e:+04d}"
    return s.rjust(w)


def _row(pid: int, width: int, values) -> str:
    return f"{int(pid):{width}d}" + "".join(_es(v) for v in values)


def reference_columns(network: str, nspec: int) -> np.ndarray:
    """for each abundance column of the reference's .pdr.fin, the
    species-file slot whose abundance it holds (netgen.reference_slot)"""
    from .netgen import reference_slot
    return np.array([reference_slot(network, c) for c in range(nspec)])


def internal_order(n: int) -> np.ndarray:
    """input indices of the reference's internal point order: the input
    order with the last two PDR points swapped (3DPDR.F90:206-232)"""
    o = np.arange(n)
    if n >= 2:
        o[[n - 2, n - 1]] = o[[n - 1, n - 2]]
    return o


def _fin_sequence(n: int) -> np.ndarray:
    """internal indices in the order .pdr.fin and .line.fin write them"""
    return np.concatenate([np.arange(n - 2), [n - 1, n - 2]]) if n >= 2 else np.arange(n)


def pdr_fin_line(pid: int, xyz, tgas: float, tdust: float, etype: int, rho: float, uv: float, abundance, av,
                 pseudo_1d: bool = False) -> str:
    if pseudo_1d:
        head = [xyz[0], av[AV_RAY], tgas, tdust]
        tail = [rho, uv, *abundance]
    else:
        head = [xyz[0], xyz[1], xyz[2], tgas, tdust]
        tail = [rho, uv, *abundance, *av]
    return f"{pid:7d}" + "".join(_es(float(v)) for v in head) + f"{int(etype):5d}" + "".join(
        _es(float(v)) for v in tail)


def write_pdr_fin(path: str, xyz, tgas, tdust, rho, uv, abundance, av, etype=None, pseudo_1d: bool = False,
                  ids=None) -> None:
    """Rows as given: xyz [n, 3] pc; tgas, tdust, rho, uv [n]; abundance
    [n, nspec]; av [n, nrays]; ids (default 1..n) the id column."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    et = np.ones(n, dtype=int) if etype is None else np.asarray(etype)
    ids = np.arange(1, n + 1) if ids is None else np.asarray(ids)
    with open(path, "w") as f:
        for p in range(n):
            f.write(pdr_fin_line(int(ids[p]), xyz[p], tgas[p], tdust[p], et[p], rho[p], uv[p], abundance[p], av[p],
                                 pseudo_1d) + "\n")


def write_reference_pdr_fin(path: str, xyz, tgas, tdust, rho, uv, abundance, av, idlist=None, etype=None,
                            pseudo_1d: bool = False) -> None:
    """.pdr.fin from arrays in the internal order, rows and ids as the
    reference writes them; idlist[pp] = input line number of internal point
    pp (default pp + 1)."""
    n = len(tgas)
    seq = _fin_sequence(n)
    idl = np.arange(1, n + 1) if idlist is None else np.asarray(idlist)
    et = np.ones(n, dtype=int) if etype is None else np.asarray(etype)
    take = lambda a: np.asarray(a)[seq]
    write_pdr_fin(path, take(np.asarray(xyz, dtype=np.float64).reshape(-1, 3)), take(tgas), take(tdust), take(rho),
                  take(uv), take(abundance), take(av), etype=take(et), pseudo_1d=pseudo_1d, ids=idl[seq])


def write_cool_fin(path: str, xyz, av, cool, idlist=None, pseudo_1d: bool = False) -> None:
    """cool [n, 5]: CII, CI, OI, CO and total, internal order."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    idl = np.arange(1, n + 1) if idlist is None else np.asarray(idlist)
    with open(path, "w") as f:
        for p in range(n):
            pos = [xyz[p, 0], av[p][AV_RAY]] if pseudo_1d else list(xyz[p])
            tail = [] if pseudo_1d else list(av[p])
            f.write(_row(idl[p], 7, pos + list(cool[p][:5]) + tail) + "\n")


def write_heat_fin(path: str, xyz, av, heat, idlist=None, pseudo_1d: bool = False) -> None:
    """heat [n, 12] (pdrchem_heating_device's columns), internal order."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    idl = np.arange(1, n + 1) if idlist is None else np.asarray(idlist)
    with open(path, "w") as f:
        for p in range(n):
            pos = [xyz[p, 0], av[p][AV_RAY]] if pseudo_1d else list(xyz[p])
            tail = [] if pseudo_1d else list(av[p])
            f.write(_row(idl[p], 7, pos + list(heat[p][:12]) + tail) + "\n")


def write_line_fin(path: str, xyz, av, lines, pseudo_1d: bool = False) -> None:
    """lines [n, 17], internal order; rows 1..n-2, n, n-1 with their index."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    with open(path, "w") as f:
        for p in _fin_sequence(n):
            pos = [xyz[p, 0], av[p][AV_RAY]] if pseudo_1d else list(xyz[p])
            tail = [] if pseudo_1d else list(av[p])
            f.write(_row(p + 1, 9, pos + list(lines[p][:NLINES]) + tail) + "\n")


def write_spop_fin(path: str, av, populations) -> None:
    """populations [n, nlevtot] (CII, CI, OI, CO levels), internal order."""
    with open(path, "w") as f:
        for p in range(len(populations)):
            f.write(_row(p + 1, 9, [av[p][AV_RAY], *populations[p]]) + "\n")


def write_opdp_fin(path: str, xyz, tau) -> None:
    """tau [n, 17]: the lines' optical depths along ray 6, internal order."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    with open(path, "w") as f:
        for p in range(xyz.shape[0]):
            f.write(_row(p + 1, 9, [xyz[p, 0], *tau[p][:NLINES]]) + "\n")


def write_cloud_pdr_fin(path: str, cloud, xyz, pseudo_1d: bool = False) -> None:
    """A grid or point cloud's current state (pdrchem.cloud.Cloud) as
    <output>.pdr.fin, rows in the cloud's order with ids 1..n; the dust
    temperature is the cloud's own when its thermal balance is on."""
    from . import cloud as CL
    n = cloud.npts
    thermal = getattr(cloud, "nlevtot", 0) > 0
    T = cloud.get(CL.PREV_TEMPERATURE if thermal else CL.TEMPERATURE)
    Td = cloud.get(CL.DUST_TEMPERATURE) if thermal else np.full(n, cloud.dust_temperature)
    write_pdr_fin(path, xyz, T, Td, cloud.density, cloud.get(CL.UVFIELD), cloud.get(CL.ABUNDANCE), cloud.get(CL.AV),
                  pseudo_1d=pseudo_1d)


def write_model_outputs(prefix: str, cloud, xyz, idlist=None, etype=None, pseudo_1d: bool = False,
                        reference_slots: bool = True) -> list:
    """Every output file of a finished run, from a cloud whose points are in
    the reference's internal order (xyz [n, 3] in that order): .pdr.fin,
    .cool.fin, .heat.fin, .line.fin, .spop.fin and, with pseudo_1d,
    .opdp.fin.  The cloud's thermal balance must be on.  reference_slots:
    abundance columns in the reference's order (``reference_columns``), else
    in species-file order.  Returns the paths."""
    from . import cloud as CL
    av = cloud.get(CL.AV)
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    ab = cloud.get(CL.ABUNDANCE)
    if reference_slots:
        ab = ab[:, reference_columns(cloud.net.name, ab.shape[1])]
    out = [prefix + s for s in (".pdr.fin", ".cool.fin", ".heat.fin", ".line.fin", ".spop.fin")]
    write_reference_pdr_fin(out[0], xyz, cloud.get(CL.PREV_TEMPERATURE), cloud.get(CL.DUST_TEMPERATURE),
                            cloud.density, cloud.get(CL.UVFIELD), ab, av, idlist=idlist, etype=etype,
                            pseudo_1d=pseudo_1d)
    write_cool_fin(out[1], xyz, av, cloud.get(CL.COOLING), idlist=idlist, pseudo_1d=pseudo_1d)
    write_heat_fin(out[2], xyz, av, cloud.get(CL.HEATING), idlist=idlist, pseudo_1d=pseudo_1d)
    write_line_fin(out[3], xyz, av, cloud.get(CL.LINES), pseudo_1d=pseudo_1d)
    write_spop_fin(out[4], av, cloud.get(CL.POPULATIONS))
    if pseudo_1d:
        out.append(prefix + ".opdp.fin")
        write_opdp_fin(out[-1], xyz, cloud.get(CL.OPTICAL_DEPTHS))
    return out


def read_fin(path: str) -> np.ndarray:
    """a .fin file as a float array (the id column included)"""
    return np.loadtxt(path, ndmin=2)
