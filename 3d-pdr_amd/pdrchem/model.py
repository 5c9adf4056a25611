"""A whole 3D-PDR run on the device: the reference's main program
(3DPDR.F90) as a driver of the library's cloud.

``read_params`` reads a params.dat (input_parameters.F90:19-58); ``run``
follows the main program with the reference makefile's options (REDUCED or
the full network, THERMALBALANCE, GUESS_TEMP, TEMP_FIX, CO_FIX, DUST, and
PSEUDO_1D when asked):

1. the points file: x, y, z (pc) and n_H per line (3DPDR.F90:166-199); the
   data of the last two PDR points swapped (:206-232);
2. initial abundances from the species file, non-H/He species scaled by the
   metallicity (:291-300); relative/absolute tolerances 1e-8 / 1e-30
   (:460-461);
3. evaluation points, Av and the field, the GUESS_TEMP gas temperature and
   the dust temperatures (:388-435, 1625-1716) -- ``Cloud.from_points`` and
   ``Cloud.thermal_init``;
4. the first chemistry iterations (:474-520) and LTE populations;
5. the thermal-balance iterations until every point converges or ITERTOT
   (:590-1241) -- ``Cloud.thermal_iterate``;
6. the output files (:1270-1500) -- ``outputs.write_model_outputs``.

Ionised (n_H <= rho_min) and dark-molecular (n_H > rho_max) elements, which
the reference writes to .ion.fin / .mol.fin and treats apart, are not
supported: ``run`` refuses an input that has them.
"""
from __future__ import annotations

import dataclasses
import os
import time
from typing import Callable, Optional

import numpy as np

from . import cloud as CL
from . import lamda, lib, outputs
from .netgen import load_network


@dataclasses.dataclass
class Params:
    input: str = "1Dn30.dat"
    output: str = "V1"
    level: int = 0
    theta_crit: float = 1.3
    chemiterations: int = 8
    itertot: int = 6000
    rho_min: float = 0.9
    rho_max: float = 1.0e12
    zeta: float = 5.0e-17          # s^-1
    v_turb: float = 1.0            # km/s
    dust_temperature: float = 20.0
    end_time: float = 1.0e7        # yr
    g2d: float = 100.0
    metallicity: float = 1.0
    omega: float = 0.42
    grain_radius: float = 1.0e-7
    co_file: str = "12co.dat"
    cii_file: str = "12c+.dat"
    ci_file: str = "12c.dat"
    oi_file: str = "16o.dat"
    Tguess: float = 10.0
    Tlow0: float = 10.0
    Thigh0: float = 8000.0
    Tmin: float = 2.7
    Tmax: float = 30000.0
    Fcrit: float = 0.005
    Tdiff: float = 0.01
    fieldchoice: str = "UNI"
    Gext: float = 10.0


_PARAM_LINES = (  # (line after the three-line headers, field) in params.dat order
    "input", "output", "level", "theta_crit", "chemiterations", "itertot", "rho_min", "rho_max", "zeta", "v_turb",
    "dust_temperature", "end_time", "g2d", "metallicity", "omega", "grain_radius", None, None, None, "co_file",
    "cii_file", "ci_file", "oi_file", None, None, None, "Tguess", "Tlow0", "Thigh0", "Tmin", "Tmax", "Fcrit",
    "Tdiff", None, None, None, "fieldchoice", "Gext")


def read_params(path: str) -> Params:
    """params.dat as input_parameters.F90:19-58 reads it: three header lines,
    then one value per line (the first item of a list-directed READ)."""
    lines = open(path).read().splitlines()[3:]
    p = Params()
    types = {f.name: f.type for f in dataclasses.fields(Params)}
    for line, name in zip(lines, _PARAM_LINES):
        if name is None:
            continue
        tok = line.split()[0]
        t = types[name]
        if t in (int, "int"):
            val = int(tok)
        elif t in (float, "float"):
            val = float(tok.replace("D", "E").replace("d", "e"))
        else:
            val = tok
        setattr(p, name, val)
    return p


def initial_abundances(net, metallicity: float, n: int) -> np.ndarray:
    """READ_SPECIES's abundances, all but H2, H and He times the metallicity
    (3DPDR.F90:291-300)"""
    x = np.array([s.abundance if s.name in ("H2", "H", "He") else s.abundance * metallicity for s in net.species])
    return np.tile(x, (n, 1))


def read_points(path: str) -> np.ndarray:
    """the points file: one x, y, z (pc), n_H per line"""
    return np.loadtxt(path, ndmin=2)[:, :4]


def run(params: Params, directory: str = ".", network: str = "reduced", pseudo_1d: bool = False,
        precision: int = lib.MIXED, coolants=None, order: int = lib.ORDER_REFERENCE, out_dir: Optional[str] = None,
        log: Optional[Callable[[str], None]] = print, write: bool = True, reference_slots: bool = True) -> dict:
    """One run.  Files (the points file, coolant files) are read from
    `directory`; outputs go to `out_dir` (default `directory`).  coolants:
    four lamda.Coolant (default: read the params' coolant files from
    `directory` when they are there, else the stored tables).
    reference_slots (default): the heating reads, and the .pdr.fin columns
    hold, the species where the reference's arrays hold them
    (netgen.reference_slot), so that a run reproduces the reference's; False
    reads and writes every species under its own label.  Returns the cloud,
    the iteration counts and the timings."""
    say = log or (lambda s: None)
    if pseudo_1d and params.level != 0:
        raise ValueError("HEALPix level must be 0 in PSEUDO_1D mode")
    pts = read_points(os.path.join(directory, params.input))
    rho_in = pts[:, 3]
    if np.any(rho_in <= params.rho_min) or np.any(rho_in > params.rho_max):
        raise NotImplementedError("ionised (n_H <= rho_min) or dark-molecular (n_H > rho_max) elements")
    n = len(pts)
    order_int = outputs.internal_order(n)
    xyz, rho = pts[order_int, :3], rho_in[order_int]
    idlist = np.arange(1, n + 1)
    net = load_network(network)
    x0 = initial_abundances(net, params.metallicity, n)
    phys = lib.default_phys(zeta=params.zeta / 1.3e-17, v_turb=params.v_turb * 1.0e5, omega=params.omega,
                            grain_radius=params.grain_radius, metallicity=params.metallicity, g2d=params.g2d)
    if coolants is None:
        files = {"CII": params.cii_file, "CI": params.ci_file, "OI": params.oi_file, "CO": params.co_file}
        if all(os.path.exists(os.path.join(directory, f)) for f in files.values()):
            coolants = lamda.read_coolants(directory, files)
        else:
            coolants = lamda.load_coolants()
    lib.set_globals(rtol=1.0e-8, atol=1.0e-30)
    t0 = time.perf_counter()
    field = CL.UNI if params.fieldchoice.strip().upper() == "UNI" else CL.ISO
    c = CL.Cloud.from_points(net, xyz, rho, gext=(params.Gext, 0.0, 0.0), fieldchoice=field, level=params.level,
                             theta_crit=params.theta_crit, dust_temperature=params.dust_temperature, phys=phys,
                             order=order)
    c.thermal_init(coolants, lib.default_thermal_params(Tmin=params.Tmin, Tmax=params.Tmax, Fcrit=params.Fcrit,
                                                        Tdiff=params.Tdiff, pseudo_ray=6 if pseudo_1d else -1),
                   reference_slots=reference_slots)
    c.set_abundances(x0)
    t1 = time.perf_counter()
    say(f"{n} PDR points, {c.nrays} rays, network {network}; set-up {t1 - t0:.2f} s")
    for ii in range(params.chemiterations):
        st = c.iterate(t_end=params.end_time, conv_tol=0.0, precision=precision)
        say(f"chemical iteration {ii + 1}: {st['solved']} solved, {st['failed']} failed")
    t2 = time.perf_counter()
    it_stats = []
    done = False
    for it in range(params.itertot):
        st = c.thermal_iterate(t_end=params.end_time, precision=precision)
        it_stats.append(st)
        say(f"iteration {st['iteration']}: levpop {st['levpop_iteration']}, "
            f"levels {st.get('level_converged', 0)}/{n}, thermal {st['thermally_converged']}/{n}"
            + (" [chemistry]" if not st["lvg"] else ""))
        if st["done"]:
            done = True
            break
    t3 = time.perf_counter()
    say(f"{'converged' if done else 'not converged'} after {len(it_stats)} iterations; "
        f"chemistry {t2 - t1:.2f} s, thermal loop {t3 - t2:.2f} s")
    files = []
    if write:
        files = outputs.write_model_outputs(os.path.join(out_dir or directory, params.output), c, xyz, idlist=idlist,
                                            pseudo_1d=pseudo_1d, reference_slots=reference_slots)
    return dict(cloud=c, xyz=xyz, idlist=idlist, iterations=len(it_stats), converged=done, stats=it_stats,
                files=files, seconds=dict(setup=t1 - t0, chemistry=t2 - t1, thermal=t3 - t2))
