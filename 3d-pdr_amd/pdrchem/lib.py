"""ctypes binding of libpdrchem.so (declared in include/pdrchem.h).

The product path: every call here goes through the HIP library.  If the
library is missing or cannot be loaded this module raises -- there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("PDRCHEM_LIB", os.path.join(PKG_DIR, "libpdrchem.so"))

FP64 = 0   # PDRCHEM_FP64
MIXED = 1  # PDRCHEM_MIXED

NETWORKS = {"reduced": (33, 329), "full": (215, 2926), "mynetwork": (58, 634)}

# every function and global declared in include/pdrchem.h
EXPORTED_FUNCTIONS = [
    "calculate_abundances_", "pdrchem_network_id", "pdrchem_solve_device", "pdrchem_set_precision",
    "pdrchem_rates_device", "pdrchem_rates_device_zeta", "pdrchem_kernel_info", "pdrchem_set_max_blocks", "pdrchem_last_error",
    "pdrchem_server_create", "pdrchem_server_submit", "pdrchem_server_progress", "pdrchem_server_poll",
    "pdrchem_server_last_error", "pdrchem_server_destroy", "pdrchem_server_graph_stats",
    "pdrchem_server_submit_device", "pdrchem_server_max_batch", "pdrchem_server_set_max_batch",
    "pdrchem_server_set_refill", "pdrchem_server_refill_stats",
    "pdrchem_cloud_set_server",
    "pdrchem_cloud_create_points", "pdrchem_trace_points", "pdrchem_cloud_set_zeta",
    "pdrchem_cloud_create", "pdrchem_cloud_set_abundances", "pdrchem_cloud_iterate", "pdrchem_cloud_get",
    "pdrchem_cloud_num_evaluation_points", "pdrchem_cloud_evaluation_points", "pdrchem_cloud_last_error",
    "pdrchem_cloud_destroy", "pdrchem_autotune", "pdrchem_tuned_config",
    "pdrchem_set_tuned_config", "pdrchem_clear_tuning", "pdrchem_heating_device",
    "pdrchem_heating_device_zeta", "pdrchem_set_mxstep", "pdrchem_mxstep", "pdrchem_set_electron_policy",
    "pdrchem_electron_policy", "pdrchem_jacobian_device", "pdrchem_solve_device_noalloc",
    "pdrchem_coolant_tables_create", "pdrchem_coolant_tables_destroy", "pdrchem_coolant_tables_levels",
    "pdrchem_level_populations_device", "pdrchem_lte_populations_device", "pdrchem_dust_temperature_device",
    "pdrchem_cloud_thermal_init", "pdrchem_cloud_thermal_iterate",
]
EXPORTED_GLOBALS = [
    ("chemistry_module_mp_relative_abundance_tolerance_", C.c_double),
    ("chemistry_module_mp_absolute_abundance_tolerance_", C.c_double),
    ("maincode_module_mp_start_time_", C.c_double),
    ("maincode_module_mp_end_time_", C.c_double),
    ("global_module_mp_nelect_", C.c_int),
]


class Phys(C.Structure):
    _fields_ = [("zeta", C.c_double), ("v_turb", C.c_double), ("omega", C.c_double),
                ("grain_radius", C.c_double), ("metallicity", C.c_double), ("g2d", C.c_double)]


def default_phys(**kw) -> Phys:
    """params.dat defaults (params.dat:11-18; zeta normalised as
    input_parameters.F90:31, v_turb in cm/s as 3DPDR.F90:47)."""
    d = dict(zeta=5.0e-17 / 1.3e-17, v_turb=1.0e5, omega=0.42, grain_radius=1.0e-7,
             metallicity=1.0, g2d=100.0)
    d.update(kw)
    return Phys(**d)


class CloudStats(C.Structure):
    _fields_ = [("npts", C.c_int), ("solved", C.c_int), ("converged", C.c_int), ("failed", C.c_int),
                ("negative_rates", C.c_int), ("rates_ms", C.c_double), ("solve_ms", C.c_double),
                ("nst", C.c_longlong), ("nfe", C.c_longlong), ("nje", C.c_longlong), ("nsetups", C.c_longlong),
                ("nni", C.c_longlong)]


class CoolantStruct(C.Structure):
    """pdrchem_coolant: one coolant's tables (pointers into arrays the caller
    keeps alive for the call)"""
    _fields_ = [("nlev", C.c_int), ("ntemp", C.c_int)] + \
        [(k, C.POINTER(C.c_double)) for k in ("energies", "weights", "A", "B", "freq", "temperatures", "col")]


class ThermalParams(C.Structure):
    _fields_ = [("Tmin", C.c_double), ("Tmax", C.c_double), ("Fcrit", C.c_double), ("Tdiff", C.c_double),
                ("pseudo_ray", C.c_int), ("dust", C.c_int), ("co_fix", C.c_int), ("temp_fix", C.c_int)]


def default_thermal_params(**kw) -> ThermalParams:
    """params.dat's thermal-balance values (params.dat:29-33) and the
    reference makefile's build options (PSEUDO_1D along ray 6, DUST,
    CO_FIX, TEMP_FIX)."""
    d = dict(Tmin=2.7, Tmax=30000.0, Fcrit=0.005, Tdiff=0.01, pseudo_ray=6, dust=1, co_fix=1, temp_fix=1)
    d.update(kw)
    return ThermalParams(**d)


class ThermalStats(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("iteration", "levpop_iteration", "level_conv", "first_time", "done", "lvg",
                                      "thermal_points", "thermally_converged", "level_converged")] + \
        [("coolant_converged", C.c_int * 4), ("chemistry_solves", C.c_int), ("failed", C.c_int)] + \
        [(k, C.c_longlong) for k in ("nst", "nfe", "nje", "nsetups", "nni")] + [("ms", C.c_double)]


class PdrChemError(RuntimeError):
    pass


_lib = None


def load(path: Optional[str] = None):
    """Load (once) and return the ctypes handle of libpdrchem.so."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise PdrChemError(f"{p} not built: run __graft_entry__.build() or `make -C 3d-pdr_amd`")
    L = C.CDLL(p, mode=C.RTLD_GLOBAL)
    dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
    L.calculate_abundances_.argtypes = [dp, dp, dp, dp, ip, ip, ip]
    L.calculate_abundances_.restype = C.c_int
    L.pdrchem_network_id.argtypes = [C.c_int, C.c_int]
    L.pdrchem_solve_device.argtypes = [C.c_int, vp, vp, vp, C.c_int, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_int, vp, vp, vp, C.c_int, vp]
    L.pdrchem_rates_device.argtypes = [C.c_int, C.POINTER(Phys), vp, vp, C.c_int, vp, vp, vp, C.c_int, vp, vp,
                                       vp]
    L.pdrchem_rates_device_zeta.argtypes = [C.c_int, C.POINTER(Phys), vp, vp, C.c_int, vp, vp, vp, vp, C.c_int,
                                            vp, vp, vp]
    L.pdrchem_kernel_info.argtypes = [C.c_int, C.c_int, ip, ip, ip, ip]
    L.pdrchem_heating_device.argtypes = [C.c_int, C.POINTER(Phys), ip, C.c_int, vp, vp, vp, vp, vp, vp, vp,
                                         vp]
    L.pdrchem_heating_device_zeta.argtypes = [C.c_int, C.POINTER(Phys), ip, C.c_int, vp, vp, vp, vp, vp, vp,
                                              vp, vp, vp]
    L.pdrchem_set_mxstep.argtypes = [C.c_longlong]
    L.pdrchem_mxstep.restype = C.c_longlong
    L.pdrchem_set_electron_policy.argtypes = [C.c_int]
    L.pdrchem_jacobian_device.argtypes = [C.c_int, C.c_int, vp, vp, vp, C.c_int, vp, vp]
    L.pdrchem_solve_device_noalloc.argtypes = [C.c_int, vp, vp, vp, C.c_int, C.c_double, C.c_double, C.c_double,
                                               C.c_double, C.c_int, vp, vp, C.c_int, vp]
    L.pdrchem_coolant_tables_create.argtypes = [C.POINTER(CoolantStruct), ip]
    L.pdrchem_coolant_tables_create.restype = vp
    L.pdrchem_coolant_tables_destroy.argtypes = [vp]
    L.pdrchem_coolant_tables_destroy.restype = None
    L.pdrchem_coolant_tables_levels.argtypes = [vp]
    L.pdrchem_level_populations_device.argtypes = [vp, C.c_int, C.POINTER(Phys), C.c_int, C.c_int, C.c_int, C.c_int,
                                                   vp, vp, vp, vp, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int,
                                                   vp, vp, vp, vp, vp, C.c_int, vp, vp]
    L.pdrchem_lte_populations_device.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp]
    L.pdrchem_dust_temperature_device.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp]
    L.pdrchem_cloud_thermal_init.argtypes = [vp, C.POINTER(CoolantStruct), ip, ip, C.POINTER(ThermalParams)]
    L.pdrchem_cloud_thermal_iterate.argtypes = [vp, C.c_double, C.c_int, C.POINTER(ThermalStats)]
    L.pdrchem_set_max_blocks.argtypes = [C.c_int]
    L.pdrchem_set_max_blocks.restype = None
    L.pdrchem_set_precision.argtypes = [C.c_int]
    L.pdrchem_set_precision.restype = None
    L.pdrchem_last_error.restype = C.c_char_p
    L.pdrchem_server_create.argtypes = [C.c_size_t, C.c_int, C.c_int]
    L.pdrchem_server_create.restype = vp
    L.pdrchem_server_submit.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, C.c_double]
    L.pdrchem_server_progress.argtypes = [vp, C.c_int]
    L.pdrchem_server_poll.argtypes = [vp, C.c_int, ip, ip, dp]
    L.pdrchem_server_last_error.argtypes = [vp]
    L.pdrchem_server_last_error.restype = C.c_char_p
    L.pdrchem_server_destroy.argtypes = [vp]
    L.pdrchem_server_destroy.restype = None
    L.pdrchem_server_graph_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.pdrchem_server_submit_device.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, C.c_double, vp, vp]
    L.pdrchem_server_max_batch.argtypes = [vp]
    L.pdrchem_server_set_max_batch.argtypes = [vp, C.c_int]
    L.pdrchem_server_set_refill.argtypes = [vp, C.c_int, C.c_double, C.c_double]
    L.pdrchem_server_refill_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_int)]
    L.pdrchem_cloud_set_server.argtypes = [vp, vp]
    L.pdrchem_cloud_set_zeta.argtypes = [vp, dp]
    L.pdrchem_cloud_create_points.argtypes = [C.c_int, C.c_int, dp, ip, C.c_int, dp, C.c_int, C.c_double, C.c_int,
                                              dp, C.c_double, C.POINTER(Phys), ip, C.c_int]
    L.pdrchem_cloud_create_points.restype = vp
    L.pdrchem_trace_points.argtypes = [C.c_int, dp, ip, C.c_int, C.c_double, C.c_int, ip,
                                       C.POINTER(C.c_longlong), ip, dp, C.c_longlong]
    L.pdrchem_trace_points.restype = C.c_longlong
    L.pdrchem_cloud_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, dp, C.c_int, C.c_double,
                                       C.c_int, dp, C.c_double, C.POINTER(Phys), ip, C.c_int]
    L.pdrchem_cloud_create.restype = vp
    L.pdrchem_cloud_set_abundances.argtypes = [vp, dp]
    L.pdrchem_cloud_iterate.argtypes = [vp, C.c_double, C.c_double, C.c_int, C.POINTER(CloudStats)]
    L.pdrchem_cloud_get.argtypes = [vp, C.c_int, dp]
    L.pdrchem_cloud_num_evaluation_points.argtypes = [vp]
    L.pdrchem_cloud_num_evaluation_points.restype = C.c_longlong
    L.pdrchem_cloud_evaluation_points.argtypes = [vp, ip, C.POINTER(C.c_longlong), ip, dp]
    L.pdrchem_cloud_last_error.argtypes = [vp]
    L.pdrchem_cloud_last_error.restype = C.c_char_p
    L.pdrchem_autotune.argtypes = [C.c_int, C.c_int, vp, vp, vp, C.c_int, C.c_double, C.c_int, vp, ip, ip, dp]
    L.pdrchem_tuned_config.argtypes = [C.c_int, C.c_int, C.c_int, ip, ip]
    L.pdrchem_set_tuned_config.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.pdrchem_clear_tuning.argtypes = []
    L.pdrchem_clear_tuning.restype = None
    L.pdrchem_cloud_destroy.argtypes = [vp]
    L.pdrchem_cloud_destroy.restype = None
    _lib = L
    return L


def last_error() -> str:
    return load().pdrchem_last_error().decode()


def set_globals(rtol=None, atol=None, start_time=None, end_time=None, nelect=None):
    """Write the module globals calculate_abundances_ reads (the role of the
    Fortran bind(c) variables of modules.F90:165-166, 383, 529-530)."""
    L = load()
    vals = dict(chemistry_module_mp_relative_abundance_tolerance_=rtol,
                chemistry_module_mp_absolute_abundance_tolerance_=atol,
                maincode_module_mp_start_time_=start_time, maincode_module_mp_end_time_=end_time,
                global_module_mp_nelect_=nelect)
    for name, ct in EXPORTED_GLOBALS:
        v = vals[name]
        if v is not None:
            ct.in_dll(L, name).value = v


def get_globals() -> dict:
    L = load()
    return {name: ct.in_dll(L, name).value for name, ct in EXPORTED_GLOBALS}


def calculate_abundances(abundance: np.ndarray, rate: np.ndarray, density: np.ndarray,
                         temperature: np.ndarray, precision: Optional[int] = None) -> int:
    """The reference call ``CALL CALCULATE_ABUNDANCES(ABUNDANCE, RATE, DENSITY,
    TEMPERATURE, NPART, NSPEC, NREAC)`` (3DPDR.F90:498): abundance
    [npart, nspec] is overwritten in place; returns the status."""
    L = load()
    for a in (abundance, rate, density, temperature):
        if a.dtype != np.float64 or not a.flags.c_contiguous:
            raise PdrChemError("arrays must be C-contiguous float64")
    npart, nspec = abundance.shape
    nreac = rate.shape[1]
    if rate.shape[0] != npart or density.shape[0] != npart or temperature.shape[0] != npart:
        raise PdrChemError("inconsistent point counts")
    if precision is not None:
        L.pdrchem_set_precision(precision)
    dp = C.POINTER(C.c_double)
    n, s, r = C.c_int(npart), C.c_int(nspec), C.c_int(nreac)
    return L.calculate_abundances_(abundance.ctypes.data_as(dp), rate.ctypes.data_as(dp),
                                   density.ctypes.data_as(dp), temperature.ctypes.data_as(dp),
                                   C.byref(n), C.byref(s), C.byref(r))


def network_id(name_or_shape) -> int:
    if isinstance(name_or_shape, str):
        nspec, nreac = NETWORKS[name_or_shape]
    else:
        nspec, nreac = name_or_shape
    return load().pdrchem_network_id(nspec, nreac)


def kernel_info(net: int, precision: int = MIXED) -> dict:
    regs, lds, scr, bpc = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    if load().pdrchem_kernel_info(net, precision, C.byref(regs), C.byref(lds), C.byref(scr), C.byref(bpc)):
        raise PdrChemError(last_error())
    return dict(num_regs=regs.value, lds_bytes=lds.value, scratch_bytes=scr.value, blocks_per_cu=bpc.value)


# ---------------------------------------------------------------------------
# device-resident interface over torch tensors (torch provides device memory
# and streams; the kernels are the library's)
# ---------------------------------------------------------------------------

def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def solve_device(net: int, abundance, rate, density, rtol=1e-8, atol=1e-30, t_start=0.0, t_end=1e7,
                 nelect=None, status=None, stats=None, counter=None, precision=MIXED, stream=None):
    """Queue a batched solve on torch device tensors (abundance [n, nspec]
    float64 in/out, rate [n, nreac], density [n]).  Returns (status, stats)."""
    import torch
    L = load()
    n = abundance.shape[0]
    dev = abundance.device
    if status is None:
        status = torch.empty(n, dtype=torch.int32, device=dev)
    if counter is None:
        counter = torch.zeros(1, dtype=torch.int32, device=dev)
    else:
        counter.zero_()
    if nelect is None:
        nelect = abundance.shape[1]
    for t in (abundance, rate, density):
        if t.dtype != torch.float64 or not t.is_contiguous():
            raise PdrChemError("device tensors must be contiguous float64")
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    rc = L.pdrchem_solve_device(net, _ptr(abundance), _ptr(rate), _ptr(density), n, rtol, atol, t_start, t_end,
                                nelect, _ptr(status), _ptr(stats), _ptr(counter), precision,
                                C.c_void_p(s.cuda_stream))
    if rc != 0:
        raise PdrChemError(last_error())
    return status, stats


def solve_device_noalloc(net: int, abundance, rate, density, rtol=1e-8, atol=1e-30, t_start=0.0, t_end=1e7,
                         nelect=None, status=None, stats=None, precision=MIXED, stream=None):
    """As solve_device through the allocation-free entry point (the library's
    fixed scratch array instead of a per-stream workspace; no counter)."""
    import torch
    L = load()
    n = abundance.shape[0]
    if status is None:
        status = torch.empty(n, dtype=torch.int32, device=abundance.device)
    if nelect is None:
        nelect = abundance.shape[1]
    for t in (abundance, rate, density):
        if t.dtype != torch.float64 or not t.is_contiguous():
            raise PdrChemError("device tensors must be contiguous float64")
    s = stream if stream is not None else torch.cuda.current_stream(abundance.device)
    rc = L.pdrchem_solve_device_noalloc(net, _ptr(abundance), _ptr(rate), _ptr(density), n, rtol, atol, t_start,
                                        t_end, nelect, _ptr(status), _ptr(stats), precision,
                                        C.c_void_p(s.cuda_stream))
    if rc != 0:
        raise PdrChemError(last_error())
    return status, stats


def autotune(net: int, abundance, rate, density, t_end=1e7, nelect=None, precision=MIXED, stream=None) -> dict:
    """Time every candidate launch shape on this batch (torch device tensors,
    as solve_device; abundance is left unchanged) and record the fastest for
    the batch's size bucket.  Returns the chosen shape and its time."""
    import torch
    L = load()
    n = abundance.shape[0]
    if nelect is None:
        nelect = abundance.shape[1]
    s = stream if stream is not None else torch.cuda.current_stream(abundance.device)
    spb, bpc, ms = C.c_int(), C.c_int(), C.c_double()
    rc = L.pdrchem_autotune(net, precision, _ptr(abundance), _ptr(rate), _ptr(density), n, t_end, nelect,
                            C.c_void_p(s.cuda_stream), C.byref(spb), C.byref(bpc), C.byref(ms))
    if rc != 0:
        raise PdrChemError("pdrchem_autotune failed")
    return {"systems_per_block": spb.value, "blocks_per_cu": bpc.value, "ms": ms.value}


def tuned_config(net: int, npts: int, precision=MIXED):
    spb, bpc = C.c_int(), C.c_int()
    r = load().pdrchem_tuned_config(net, precision, npts, C.byref(spb), C.byref(bpc))
    return (spb.value, bpc.value) if r == 1 else None


def clear_tuning():
    load().pdrchem_clear_tuning()


ORDER_REFERENCE = 0   # PDRCHEM_ORDER_REFERENCE: heapsort.F90's candidate order
ORDER_DISTANCE = 1    # PDRCHEM_ORDER_DISTANCE: (distance, point index)


def trace_points(xyz, level=0, theta_crit=1.3, order=ORDER_DISTANCE, etype=None):
    """Evaluation points of a point cloud (the library's host tracer; no
    device use).  Returns (count[n, nrays], offsets[n*nrays+1], proj, step)."""
    L = load()
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    nrays = 12 << (2 * level)
    et = None if etype is None else np.ascontiguousarray(etype, dtype=np.int32)
    etp = None if et is None else et.ctypes.data_as(C.POINTER(C.c_int))
    count = np.zeros((n, nrays), dtype=np.int32)
    offs = np.zeros(n * nrays + 1, dtype=np.int64)
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    tot = L.pdrchem_trace_points(n, xyz.ctypes.data_as(dp), etp, level, theta_crit, order, count.ctypes.data_as(ip),
                                 offs.ctypes.data_as(C.POINTER(C.c_longlong)), None, None, 0)
    if tot < 0:
        raise PdrChemError("bad argument to pdrchem_trace_points")
    proj = np.zeros(max(tot, 1), dtype=np.int32)
    step = np.zeros(max(tot, 1))
    L.pdrchem_trace_points(n, xyz.ctypes.data_as(dp), etp, level, theta_crit, order, count.ctypes.data_as(ip),
                           offs.ctypes.data_as(C.POINTER(C.c_longlong)), proj.ctypes.data_as(ip),
                           step.ctypes.data_as(dp), tot)
    return count, offs, proj[:tot], step[:tot]


def set_tuned_config(net: int, npts: int, systems_per_block: int, blocks_per_cu: int, precision=MIXED):
    if load().pdrchem_set_tuned_config(net, precision, npts, systems_per_block, blocks_per_cu) != 0:
        raise PdrChemError("bad launch shape")


XE_LAST_RHS = 0     # PDRCHEM_XE_LAST_RHS: x_e of the last RHS evaluation (calculate_abundances.c:317)
XE_CHARGE_SUM = 1   # PDRCHEM_XE_CHARGE_SUM: charge sum of the abundances written back


def set_mxstep(mxstep: int) -> None:
    if load().pdrchem_set_mxstep(int(mxstep)) != 0:
        raise PdrChemError(last_error())


def mxstep() -> int:
    return int(load().pdrchem_mxstep())


def set_electron_policy(policy: int) -> None:
    if load().pdrchem_set_electron_policy(int(policy)) != 0:
        raise PdrChemError(last_error())


def electron_policy() -> int:
    return int(load().pdrchem_electron_policy())


def jacobian_device(net: int, abundance, rate, density, precision=FP64, stream=None):
    """The solver's Jacobian at the given abundances (torch device tensors as
    for solve_device).  Returns [n, neq, neq+1]: d ydot / d x at fixed x_e
    and, in the last column, d ydot / d x_e."""
    import torch
    L = load()
    n, nspec = abundance.shape
    for t in (abundance, rate, density):
        if t.dtype != torch.float64 or not t.is_contiguous() or not t.is_cuda:
            raise PdrChemError("device tensors must be contiguous float64")
    out = torch.empty((n, nspec - 1, nspec), dtype=torch.float64, device=abundance.device)
    s = stream if stream is not None else torch.cuda.current_stream(abundance.device)
    if L.pdrchem_jacobian_device(net, precision, _ptr(abundance), _ptr(rate), _ptr(density), n, _ptr(out),
                                 C.c_void_p(s.cuda_stream)) != 0:
        raise PdrChemError(last_error())
    return out


def heating_device(net: int, index, density, T, Td, uv, abundance, rate, phys: Optional[Phys] = None, out=None,
                   stream=None, zeta=None):
    """Device heating terms (CALC_HEATING): index = the 13 1-based slots NH,
    NH2, NC, NCO, NHEx, NH3x, NH3Ox, NHCOx, NH2x, NELECT, NRGR, NRH2, NRCI;
    density, T, Td, uv [n]; abundance [n, nspec]; rate [n, nreac] (float64
    CUDA tensors).  Returns heat [n, 12]."""
    import torch
    L = load()
    args = [density, T, Td, uv, abundance, rate]
    for t in args:
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise PdrChemError("heating_device takes contiguous float64 device tensors")
    npts = density.shape[0]
    dims = dict(enumerate(NETWORKS.values())).get(net)
    if dims is None or any(t.numel() != npts for t in args[1:4]) or abundance.numel() != npts * dims[0] \
            or rate.numel() != npts * dims[1]:
        raise PdrChemError("heating_device: unsupported network or shapes that do not match it")
    if out is None:
        out = torch.empty((npts, 12), dtype=torch.float64, device=density.device)
    idx = (C.c_int * 13)(*[int(i) for i in index])
    ph = phys or default_phys()
    s = stream if stream is not None else torch.cuda.current_stream(density.device).cuda_stream
    if zeta is not None and not (zeta.is_cuda and zeta.dtype == torch.float64 and zeta.numel() == npts):
        raise PdrChemError("heating_device: zeta must be a float64 device tensor of one value per point")
    rc = L.pdrchem_heating_device_zeta(net, C.byref(ph), idx, npts, *[t.data_ptr() for t in args],
                                       zeta.data_ptr() if zeta is not None else None, out.data_ptr(), s)
    if rc != 0:
        raise PdrChemError("pdrchem_heating_device failed (network, index or arguments)")
    return out


def rates_device(net: int, T, Td, G, Av, cols=None, phys: Optional[Phys] = None, out=None, neg=None,
                 stream=None, zeta=None):
    """Device rate evaluation: T, Td [n]; G, Av [n, nrays]; cols [n, nrays, 5]
    (H2, HD, CO, C, S) or None; zeta [n] per-point cosmic-ray ionisation rate
    (units of phys.zeta) or None.  Returns (rate [n, nreac], neg counter)."""
    import torch
    L = load()
    n = T.shape[0]
    nrays = G.shape[1] if G.dim() > 1 else 1
    nreac = [v for v in NETWORKS.values()][net][1]
    if out is None:
        out = torch.empty((n, nreac), dtype=torch.float64, device=T.device)
    if neg is None:
        neg = torch.zeros(1, dtype=torch.int32, device=T.device)
    ph = phys or default_phys()
    s = stream if stream is not None else torch.cuda.current_stream(T.device)
    rc = L.pdrchem_rates_device_zeta(net, C.byref(ph), _ptr(T), _ptr(Td), nrays, _ptr(G), _ptr(Av), _ptr(cols),
                                     _ptr(zeta), n, _ptr(out), _ptr(neg), C.c_void_p(s.cuda_stream))
    if rc != 0:
        raise PdrChemError(last_error())
    return out, neg


# ---------------------------------------------------------------------------
# request loop (continuous batching) over host arrays
# ---------------------------------------------------------------------------

class Server:
    """Thin owner of a pdrchem_server: submit requests of host arrays, drive
    progress, collect (ticket, status, latency_ms).  The arrays of a request
    are kept referenced until it completes (the library writes the
    abundances back into them)."""

    def __init__(self, device_bytes: int = 1 << 30, nslots: int = 4, precision: int = MIXED):
        self.L = load()
        self.h = self.L.pdrchem_server_create(device_bytes, nslots, precision)
        if not self.h:
            raise PdrChemError("pdrchem_server_create failed")
        self.live = {}

    def submit(self, net: int, abundance: np.ndarray, rate: np.ndarray, density: np.ndarray, t_end_years: float) -> int:
        for a in (abundance, rate, density):
            if a.dtype != np.float64 or not a.flags.c_contiguous:
                raise PdrChemError("arrays must be C-contiguous float64")
        n = abundance.shape[0]
        t = self.L.pdrchem_server_submit(self.h, net, n, abundance.ctypes.data, rate.ctypes.data,
                                         density.ctypes.data, float(t_end_years))
        if t < 0:
            raise PdrChemError(self.L.pdrchem_server_last_error(self.h).decode())
        self.live[t] = (abundance, rate, density)
        return t

    def submit_device(self, net: int, abundance, rate, density, t_end_years: float, status=None, stats=None) -> int:
        """Device request: torch device tensors (abundance [n, nspec] float64,
        updated in place; rate [n, nreac]; density [n]; optional status [n]
        int32 and stats [n, 8] int32 outputs).  Kept referenced until done."""
        import torch
        for a in (abundance, rate, density):
            if not a.is_cuda or a.dtype != torch.float64 or not a.is_contiguous():
                raise PdrChemError("device requests take contiguous float64 cuda tensors")
        n = abundance.shape[0]
        t = self.L.pdrchem_server_submit_device(self.h, net, n, _ptr(abundance), _ptr(rate), _ptr(density),
                                                float(t_end_years), _ptr(status), _ptr(stats))
        if t < 0:
            raise PdrChemError(self.L.pdrchem_server_last_error(self.h).decode())
        self.live[t] = (abundance, rate, density, status, stats)
        return t

    def progress(self, wait: bool = False) -> int:
        r = self.L.pdrchem_server_progress(self.h, 1 if wait else 0)
        if r < 0:
            raise PdrChemError(self.L.pdrchem_server_last_error(self.h).decode())
        return r

    def poll(self, max_n: int = 4096):
        tk = (C.c_int * max_n)()
        st = (C.c_int * max_n)()
        lat = (C.c_double * max_n)()
        n = self.L.pdrchem_server_poll(self.h, max_n, tk, st, lat)
        out = [(tk[i], st[i], lat[i]) for i in range(n)]
        for t, _, _ in out:
            self.live.pop(t, None)
        return out

    def set_max_batch(self, npts: int) -> None:
        if self.L.pdrchem_server_set_max_batch(self.h, npts) != 0:
            raise PdrChemError("bad batch size")

    def set_refill(self, ring_points: int, idle_ms: float = 2.0, life_ms: float = 250.0) -> None:
        """Standing launches for host requests (ring_points a power of two,
        >= 64; 0: batches).  Refused while standing launches hold requests."""
        if self.L.pdrchem_server_set_refill(self.h, ring_points, idle_ms, life_ms) != 0:
            raise PdrChemError("set_refill: " + self.L.pdrchem_server_last_error(self.h).decode())

    def refill_stats(self):
        """(ring points or 0, standing launches so far, points held in the rings)."""
        la, held = C.c_longlong(), C.c_int()
        r = self.L.pdrchem_server_refill_stats(self.h, C.byref(la), C.byref(held))
        return r, la.value, held.value

    def graph_stats(self):
        """(uses_graphs, graph launches, graph instantiations)."""
        la, ins = C.c_longlong(), C.c_longlong()
        g = self.L.pdrchem_server_graph_stats(self.h, C.byref(la), C.byref(ins))
        return bool(g), la.value, ins.value

    def close(self):
        if self.h:
            self.L.pdrchem_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------------
# thermal balance: coolant tables, level populations, dust temperature
# --------------------------------------------------------------------------

THERMAL_SPECIES = ("NCx", "NC", "NO", "NCO", "NH", "NPROTON", "NELECT", "NHE", "NH2")


def thermal_slots(net) -> np.ndarray:
    """0-based abundance slots of C+, C, O, CO (the coolants) and H, H+, e-,
    He, H2 (the collision partners), from READ_SPECIES's labels"""
    return np.array([net.labels[k] - 1 for k in THERMAL_SPECIES], dtype=np.int32)


def heating_index(net, reference_slots: bool = False) -> np.ndarray:
    """the 13 1-based slots of pdrchem_heating_device: NH, NH2, NC, NCO,
    NHEx, NH3x, NH3Ox, NHCOx, NH2x, NELECT, NRGR, NRH2, NRCI.  With
    reference_slots, each species label points where the reference's own
    arrays hold it (netgen.reference_slot: its reduced network reads CH5+ as
    NH2x, sub_calculate_heating.F90:229-230)."""
    from .netgen import reference_slot
    sp = [net.labels[k] for k in ("NH", "NH2", "NC", "NCO", "NHEx", "NH3x", "NH3Ox", "NHCOx", "NH2x", "NELECT")]
    if reference_slots:
        sp = [reference_slot(net.name, s - 1) + 1 for s in sp]
    return np.array(sp + [net.nr_indices[k] for k in ("NRGR", "NRH2", "NRCI")], dtype=np.int32)


def coolant_structs(coolants):
    """(pdrchem_coolant[4], arrays to keep alive) from four lamda.Coolant"""
    if len(coolants) != 4:
        raise PdrChemError("four coolants (CII, CI, OI, CO) expected")
    arr = (CoolantStruct * 4)()
    keep = []
    dp = C.POINTER(C.c_double)
    for k, c in enumerate(coolants):
        fa = [np.ascontiguousarray(a, dtype=np.float64) for a in c.fortran_arrays()]
        keep.extend(fa)
        arr[k] = CoolantStruct(c.nlev, c.ntemp, *[a.ctypes.data_as(dp) for a in fa])
    return arr, keep


class CoolantTables:
    """Device copies of the four coolants (pdrchem_coolant_tables)."""

    def __init__(self, coolants, slots):
        L = load()
        arr, keep = coolant_structs(coolants)
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        if sl.shape != (9,):
            raise PdrChemError("9 abundance slots expected")
        self.h = L.pdrchem_coolant_tables_create(arr, sl.ctypes.data_as(C.POINTER(C.c_int)))
        del keep
        if not self.h:
            raise PdrChemError(L.pdrchem_last_error().decode())
        self.nlevtot = L.pdrchem_coolant_tables_levels(self.h)
        self.offsets = np.cumsum([0] + [c.nlev for c in coolants])

    def close(self):
        if getattr(self, "h", None):
            load().pdrchem_coolant_tables_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def level_populations_device(tables: CoolantTables, coolant: int, T, Td, rho, ab, pop, count, offs, proj, step,
                             nrays: int, pseudo_ray: int = -1, phys: Optional[Phys] = None, co_fix: int = 0,
                             co_average_co: int = 0, levpop_iteration: int = 0, skip=None, solution=None,
                             cool=None, lines=None, tau_ray: Optional[int] = None, tau_lines=None, stream=None):
    """Escape-probability level populations of one coolant (-1: all four) on
    device tensors (float64; count/proj int32, offs int64, skip uint8):
    T, Td, rho [n]; ab [n, nspec]; pop [nrows, nlevtot]; the evaluation
    points of ray j of point p are rows proj[offs[p*nrays+j] + i] of pop,
    i < count[p*nrays+j], at steps `step` (pc).  Returns (solution [n,
    nlevtot], cool [n, 4], lines [n, 17]), and the lines' optical depths
    along ray tau_ray [n, 17] as a fourth item when tau_ray is given;
    outputs may be passed in."""
    import torch
    n = T.shape[0]
    dev = T.device
    f64 = dict(dtype=torch.float64, device=dev)
    solution = solution if solution is not None else torch.zeros(n, tables.nlevtot, **f64)
    cool = cool if cool is not None else torch.zeros(n, 4, **f64)
    lines = lines if lines is not None else torch.zeros(n, 17, **f64)
    S = torch.empty(max(n * nrays * tables.nlevtot, 1), **f64)
    if tau_ray is not None and tau_lines is None:
        tau_lines = torch.zeros(n, 17, **f64)
    ph = phys or default_phys()
    s = stream if stream is not None else torch.cuda.current_stream(dev).cuda_stream
    ptr = lambda t: None if t is None else t.data_ptr()
    rc = load().pdrchem_level_populations_device(
        tables.h, coolant, C.byref(ph), n, ab.shape[1], nrays, pseudo_ray, ptr(T), ptr(Td), ptr(rho), ptr(ab),
        ptr(pop), ptr(count), ptr(offs), ptr(proj), ptr(step), co_fix, co_average_co, levpop_iteration, ptr(skip),
        S.data_ptr(), solution.data_ptr(), cool.data_ptr(), lines.data_ptr(), -1 if tau_ray is None else tau_ray,
        ptr(tau_lines), s)
    if rc != 0:
        raise PdrChemError(load().pdrchem_last_error().decode())
    return (solution, cool, lines) if tau_ray is None else (solution, cool, lines, tau_lines)


def lte_populations_device(tables: CoolantTables, T, ab, rho, skip=None, pop=None, stream=None):
    """LTE populations of every coolant [n, nlevtot] (device tensors)."""
    import torch
    n = T.shape[0]
    pop = pop if pop is not None else torch.zeros(n, tables.nlevtot, dtype=torch.float64, device=T.device)
    s = stream if stream is not None else torch.cuda.current_stream(T.device).cuda_stream
    rc = load().pdrchem_lte_populations_device(tables.h, n, ab.shape[1], T.data_ptr(), ab.data_ptr(), rho.data_ptr(),
                                               None if skip is None else skip.data_ptr(), pop.data_ptr(), s)
    if rc != 0:
        raise PdrChemError(load().pdrchem_last_error().decode())
    return pop


def dust_temperature_device(G, uv, out=None, stream=None):
    """CALCULATE_DUST_TEMPERATURES on device tensors: G [n, nrays], uv [n]."""
    import torch
    n, nrays = G.shape
    out = out if out is not None else torch.zeros(n, dtype=torch.float64, device=G.device)
    s = stream if stream is not None else torch.cuda.current_stream(G.device).cuda_stream
    if load().pdrchem_dust_temperature_device(n, nrays, G.data_ptr(), uv.data_ptr(), out.data_ptr(), s) != 0:
        raise PdrChemError(load().pdrchem_last_error().decode())
    return out
