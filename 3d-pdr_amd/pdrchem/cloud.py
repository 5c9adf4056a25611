"""Grid cloud: ray-traced column densities and the chemistry iteration, on
the device (the pdrchem_cloud_* functions of libpdrchem.so).

The reference runs these steps inside its main program: evaluation points
along the HEALPix rays (evaluation_points, eval_points.F90:129-225), Av and
the surface field per ray (CALC_UVFIELD, 3DPDR.F90:1636-1708), column
densities (calc_columndens, 3DPDR.F90:1788-1844) and the chemistry
iterations of rates, solve and new columns (3DPDR.F90:474-520).  Here a
``Cloud`` holds a regular grid on the device and runs them there; only
statistics and requested fields come back to the host.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import lib
from .netgen import Network

ABUNDANCE, AV, G, COLUMNS, TEMPERATURE, UVFIELD = range(6)
(DUST_TEMPERATURE, HEATING, COOLING, LINES, POPULATIONS, PREV_TEMPERATURE,
 THERMAL_CONVERGED, OPTICAL_DEPTHS) = range(6, 14)
SHIELD_SPECIES = ("H2", "HD", "CO", "C", "S")   # column order expected by the rate kernel
ISO, UNI = 0, 1


def shield_slots(net: Network) -> np.ndarray:
    names = [s.name for s in net.species]
    return np.array([names.index(s) if s in names else -1 for s in SHIELD_SPECIES], dtype=np.int32)


class Cloud:
    """A regular nx*ny*nz grid (point index ix + nx*(iy + ny*iz)) with
    spacing ``spacing_pc`` and density n_H per point, illuminated by
    ``gext`` (UNI: a vector in Draine units; ISO: gext[0] spread over the
    rays), HEALPix ``level`` and ``theta_crit`` as params.dat."""

    def __init__(self, net: Network, shape: Sequence[int], spacing_pc: float, density: np.ndarray,
                 gext=(100.0, 0.0, 0.0), fieldchoice: int = UNI, level: int = 0, theta_crit: float = 1.3,
                 dust_temperature: float = 20.0, phys: Optional[lib.Phys] = None):
        L = lib.load()
        self.net = net
        self.nx, self.ny, self.nz = (int(v) for v in shape)
        self.npts = self.nx * self.ny * self.nz
        self.nrays = 12 << (2 * level)
        self.nspec = net.nspec
        rho = np.ascontiguousarray(density, dtype=np.float64).reshape(-1)
        if rho.shape[0] != self.npts:
            raise lib.PdrChemError("density must have nx*ny*nz entries")
        self.density = rho.copy()
        self.dust_temperature = float(dust_temperature)
        nid = lib.network_id(net.name)
        g = np.ascontiguousarray(gext, dtype=np.float64)
        self._phys = phys or lib.default_phys()
        slots = shield_slots(net)
        dp = C.POINTER(C.c_double)
        self._h = L.pdrchem_cloud_create(nid, self.nx, self.ny, self.nz, spacing_pc, rho.ctypes.data_as(dp), level,
                                         theta_crit, fieldchoice, g.ctypes.data_as(dp), dust_temperature,
                                         C.byref(self._phys), slots.ctypes.data_as(C.POINTER(C.c_int)), net.nspec)
        if not self._h:
            raise lib.PdrChemError(L.pdrchem_cloud_last_error(None).decode())

    @classmethod
    def from_points(cls, net: Network, xyz: np.ndarray, density: np.ndarray, gext=(100.0, 0.0, 0.0),
                    fieldchoice: int = UNI, level: int = 0, theta_crit: float = 1.3, dust_temperature: float = 20.0,
                    phys: Optional[lib.Phys] = None, order: int = lib.ORDER_DISTANCE,
                    etype: Optional[np.ndarray] = None) -> "Cloud":
        """A cloud of arbitrary points (x, y, z in pc; n_H), as the reference's
        input files hold; evaluation points in `order` (ORDER_REFERENCE: the
        reference's heapsort order)."""
        L = lib.load()
        self = cls.__new__(cls)
        self.net = net
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.npts = xyz.shape[0]
        self.nx, self.ny, self.nz = self.npts, 1, 1
        self.nrays = 12 << (2 * level)
        self.nspec = net.nspec
        rho = np.ascontiguousarray(density, dtype=np.float64).reshape(-1)
        if rho.shape[0] != self.npts:
            raise lib.PdrChemError("density must have one entry per point")
        self.density = rho.copy()
        self.dust_temperature = float(dust_temperature)
        g = np.ascontiguousarray(gext, dtype=np.float64)
        self._phys = phys or lib.default_phys()
        slots = shield_slots(net)
        et = None if etype is None else np.ascontiguousarray(etype, dtype=np.int32)
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
        self._h = L.pdrchem_cloud_create_points(lib.network_id(net.name), self.npts, xyz.ctypes.data_as(dp),
                                                None if et is None else et.ctypes.data_as(ip), order,
                                                rho.ctypes.data_as(dp), level, theta_crit, fieldchoice,
                                                g.ctypes.data_as(dp), dust_temperature, C.byref(self._phys),
                                                slots.ctypes.data_as(ip), net.nspec)
        if not self._h:
            raise lib.PdrChemError(L.pdrchem_cloud_last_error(None).decode())
        return self

    def _err(self) -> str:
        return lib.load().pdrchem_cloud_last_error(self._h).decode()

    def use_server(self, server: Optional[lib.Server]) -> None:
        """Send the solves of later iterations through this request loop
        (None: direct launches)."""
        self._server = server   # keep it alive while in use
        if lib.load().pdrchem_cloud_set_server(self._h, server.h if server is not None else None) != 0:
            raise lib.PdrChemError(self._err())

    def set_zeta(self, zeta: Optional[np.ndarray]) -> None:
        """Per-point cosmic-ray ionisation rate (units of 1.3e-17 s^-1, like
        Phys.zeta) for later iterations; None: the global Phys.zeta."""
        L = lib.load()
        if zeta is None:
            rc = L.pdrchem_cloud_set_zeta(self._h, None)
        else:
            z = np.ascontiguousarray(zeta, dtype=np.float64).reshape(-1)
            if z.shape[0] != self.npts:
                raise lib.PdrChemError("zeta must have one entry per point")
            rc = L.pdrchem_cloud_set_zeta(self._h, z.ctypes.data_as(C.POINTER(C.c_double)))
        if rc != 0:
            raise lib.PdrChemError(self._err())

    def set_abundances(self, x: np.ndarray) -> None:
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (self.npts, self.nspec):
            raise lib.PdrChemError(f"abundances must be [{self.npts}, {self.nspec}]")
        if lib.load().pdrchem_cloud_set_abundances(self._h, x.ctypes.data_as(C.POINTER(C.c_double))) != 0:
            raise lib.PdrChemError(self._err())

    def iterate(self, t_end: float = 1e7, conv_tol: float = 1e-3, precision: int = lib.MIXED) -> dict:
        st = lib.CloudStats()
        if lib.load().pdrchem_cloud_iterate(self._h, t_end, conv_tol, precision, C.byref(st)) != 0:
            raise lib.PdrChemError(self._err())
        return {k: getattr(st, k) for k, _ in st._fields_}

    def thermal_init(self, coolants=None, params: Optional[lib.ThermalParams] = None,
                     reference_slots: bool = False) -> None:
        """Turn on the thermal balance (THERMALBALANCE, GUESS_TEMP): the four
        coolants (lamda.Coolant; default the stored CII, CI, OI, CO tables)
        and the parameters (default params.dat's and the reference build's).
        reference_slots: the heating reads its species where the reference's
        arrays hold them (lib.heating_index)."""
        from . import lamda
        L = lib.load()
        cools = coolants if coolants is not None else lamda.load_coolants()
        arr, keep = lib.coolant_structs(cools)
        slots = lib.thermal_slots(self.net)
        hidx = lib.heating_index(self.net, reference_slots)
        self._tparams = params or lib.default_thermal_params()
        ip = C.POINTER(C.c_int)
        rc = L.pdrchem_cloud_thermal_init(self._h, arr, slots.ctypes.data_as(ip), hidx.ctypes.data_as(ip),
                                          C.byref(self._tparams))
        del keep
        if rc != 0:
            raise lib.PdrChemError(self._err())
        self.nlevtot = sum(c.nlev for c in cools)

    def thermal_iterate(self, t_end: float = 1e7, precision: int = lib.MIXED) -> dict:
        """One iteration of the reference's outer loop (3DPDR.F90:590-1241)."""
        st = lib.ThermalStats()
        if lib.load().pdrchem_cloud_thermal_iterate(self._h, t_end, precision, C.byref(st)) != 0:
            raise lib.PdrChemError(self._err())
        out = {k: getattr(st, k) for k, _ in st._fields_}
        out["coolant_converged"] = list(st.coolant_converged)
        return out

    def get(self, field: int) -> np.ndarray:
        shapes = {ABUNDANCE: (self.npts, self.nspec), AV: (self.npts, self.nrays), G: (self.npts, self.nrays),
                  COLUMNS: (self.npts, self.nrays, 5), TEMPERATURE: (self.npts,), UVFIELD: (self.npts,),
                  DUST_TEMPERATURE: (self.npts,), HEATING: (self.npts, 12), COOLING: (self.npts, 5),
                  LINES: (self.npts, 17), POPULATIONS: (self.npts, getattr(self, "nlevtot", 0)),
                  PREV_TEMPERATURE: (self.npts,), THERMAL_CONVERGED: (self.npts,),
                  OPTICAL_DEPTHS: (self.npts, 17)}
        out = np.zeros(shapes[field])
        if lib.load().pdrchem_cloud_get(self._h, field, out.ctypes.data_as(C.POINTER(C.c_double))) != 0:
            raise lib.PdrChemError(self._err())
        return out

    def evaluation_points(self):
        """(count[npts, nrays], offsets[npts*nrays+1], proj[total], step[total])."""
        L = lib.load()
        total = L.pdrchem_cloud_num_evaluation_points(self._h)
        count = np.zeros((self.npts, self.nrays), dtype=np.int32)
        offs = np.zeros(self.npts * self.nrays + 1, dtype=np.int64)
        proj = np.zeros(max(total, 1), dtype=np.int32)
        step = np.zeros(max(total, 1))
        rc = L.pdrchem_cloud_evaluation_points(self._h, count.ctypes.data_as(C.POINTER(C.c_int)),
                                               offs.ctypes.data_as(C.POINTER(C.c_longlong)),
                                               proj.ctypes.data_as(C.POINTER(C.c_int)),
                                               step.ctypes.data_as(C.POINTER(C.c_double)))
        if rc != 0:
            raise lib.PdrChemError(self._err())
        return count, offs, proj[:total], step[:total]

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.load().pdrchem_cloud_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
