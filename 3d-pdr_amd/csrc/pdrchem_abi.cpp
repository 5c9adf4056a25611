// C ABI of libpdrchem.so (declarations and reference citations in
// include/pdrchem.h).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pdrchem.h"
#include "gen/net_reduced.h"
#include "gen/wg_full.h"
#include "gen/wg_mynetwork.h"
#include "pdr_internal.h"

using namespace pdrchem;

// ---------------------------------------------------------------------------
// Globals of the reference interface (weak: a Fortran host's bind(c)
// definitions take precedence).  Defaults as set by the reference driver:
// tolerances 3DPDR.F90:462-463, start time 3DPDR.F90:452, end time params.dat.
// ---------------------------------------------------------------------------
extern "C" {
__attribute__((weak)) double chemistry_module_mp_relative_abundance_tolerance_ = 1.0e-8;
__attribute__((weak)) double chemistry_module_mp_absolute_abundance_tolerance_ = 1.0e-30;
__attribute__((weak)) double maincode_module_mp_start_time_ = 0.0;
__attribute__((weak)) double maincode_module_mp_end_time_ = 1.0e7;
__attribute__((weak)) int global_module_mp_nelect_ = 0;
}

namespace {

constexpr double SECONDS_IN_YEAR = 3.1556926e7;  // calculate_abundances.c:90

thread_local std::string g_err;
int g_max_blocks = 0;
int g_precision = -1;  // -1: not set, read PDRCHEM_PRECISION on first use
long long g_mxstep = 10000000LL;  // calculate_abundances.c:111
int g_xe_policy = PDRCHEM_XE_LAST_RHS;

int host_precision() {
  if (g_precision < 0) {
    const char *e = std::getenv("PDRCHEM_PRECISION");
    g_precision = (e && std::strcmp(e, "fp64") == 0) ? PDRCHEM_FP64 : PDRCHEM_MIXED;
  }
  return g_precision;
}

void set_err(const std::string &s) { g_err = s; }

bool hip_ok(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  set_err(std::string(what) + ": " + hipGetErrorString(e));
  return false;
}

struct NetShape {
  int nspec, nreac;
};
const NetShape kNets[NET_COUNT] = {
    {pdrnet_reduced::NSPEC, pdrnet_reduced::NREAC},
    {pdrwg_full::NSPEC, pdrwg_full::NREAC},
    {pdrwg_mynetwork::NSPEC, pdrwg_mynetwork::NREAC},
};
const bool kCompiled[NET_COUNT] = {true, true, true};

int launch_solver(int net, const SolveRequest &r, hipStream_t s) {
  return net == NET_REDUCED ? launch_wave_solver(net, r, s) : launch_wg_solver(net, r, s);
}

int solver_info(int net, int mixed, KernelInfo *info) {
  return net == NET_REDUCED ? wave_solver_info(net, mixed, info) : wg_solver_info(net, mixed, info);
}

// device buffers reused across calculate_abundances_ calls
struct HostPathBuffers {
  std::mutex mu;
  double *ab = nullptr, *rate = nullptr, *dens = nullptr;
  int *status = nullptr;
  unsigned int *counter = nullptr;
  size_t cap_ab = 0, cap_rate = 0, cap_pts = 0;
  hipStream_t stream = nullptr;

  bool reserve(size_t n_ab, size_t n_rate, size_t npts) {
    if (!stream && !hip_ok(hipStreamCreate(&stream), "hipStreamCreate")) return false;
    if (!counter && !hip_ok(hipMalloc(&counter, sizeof(unsigned int)), "hipMalloc")) return false;
    if (n_ab > cap_ab) {
      if (ab) (void)hipFree(ab);
      ab = nullptr;
      if (!hip_ok(hipMalloc(&ab, n_ab * sizeof(double)), "hipMalloc(abundance)")) return false;
      cap_ab = n_ab;
    }
    if (n_rate > cap_rate) {
      if (rate) (void)hipFree(rate);
      rate = nullptr;
      if (!hip_ok(hipMalloc(&rate, n_rate * sizeof(double)), "hipMalloc(rate)")) return false;
      cap_rate = n_rate;
    }
    if (npts > cap_pts) {
      if (dens) (void)hipFree(dens);
      if (status) (void)hipFree(status);
      dens = nullptr;
      status = nullptr;
      if (!hip_ok(hipMalloc(&dens, npts * sizeof(double)), "hipMalloc(density)")) return false;
      if (!hip_ok(hipMalloc(&status, npts * sizeof(int)), "hipMalloc(status)")) return false;
      cap_pts = npts;
    }
    return true;
  }
};
HostPathBuffers g_buf;

// cvode.log entries for failed points, in the reference's layout
// (calculate_abundances.c:108, 235-298): the fifth-failure block for points
// that ran out of retries on -1 or -4, the plain failure block for any other
// flag.  The per-call retry lines and the time reached at the failure are not
// reported by the device, so "After" gives the flag only.
void log_failures(const std::vector<int> &st, double start_time, double end_time) {
  FILE *fh = nullptr;
  for (size_t n = 0; n < st.size(); n++) {
    if (st[n] == 0) continue;
    if (!fh) fh = std::fopen("cvode.log", "a");
    if (!fh) return;
    if (st[n] == -1 || st[n] == -4)
      std::fprintf(fh, "ERROR! Fifth failure - aborting integration\n\n");
    else
      std::fprintf(fh, "ERROR! CVODE failed with the following parameters:\n\n");
    std::fprintf(fh, "-----------------------------------------------\n");
    std::fprintf(fh, "CVODE Parameters:\n\n");
    std::fprintf(fh, "Particle %zu\n\n", n);
    std::fprintf(fh, "Before:\n");
    std::fprintf(fh, "flag   = %d\n", 0);
    std::fprintf(fh, "t      = %8.2le yr\n", start_time);
    std::fprintf(fh, "tout   = %8.2le yr\n", end_time);
    std::fprintf(fh, "dt     = %8.2le yr\n\n", end_time - start_time);
    std::fprintf(fh, "After:\n");
    std::fprintf(fh, "flag   = %d\n", st[n]);
    std::fprintf(fh, "-----------------------------------------------\n\n");
  }
  if (fh) std::fclose(fh);
}

}  // namespace

namespace pdrchem {
SolverOptions solver_options() { return SolverOptions{g_mxstep, g_xe_policy}; }
void set_last_error(const std::string &msg) { set_err(msg); }
}  // namespace pdrchem

extern "C" {

const char *pdrchem_last_error(void) { return g_err.c_str(); }

int pdrchem_set_mxstep(long long mxstep) {
  if (mxstep < 1) {
    set_err("mxstep must be positive");
    return -1;
  }
  g_mxstep = mxstep;
  return 0;
}

long long pdrchem_mxstep(void) { return g_mxstep; }

int pdrchem_set_electron_policy(int policy) {
  if (policy != PDRCHEM_XE_LAST_RHS && policy != PDRCHEM_XE_CHARGE_SUM) {
    set_err("unknown electron policy");
    return -1;
  }
  g_xe_policy = policy;
  return 0;
}

int pdrchem_electron_policy(void) { return g_xe_policy; }

int pdrchem_jacobian_device(int net, int precision, const double *d_abundance, const double *d_rate,
                            const double *d_density, int npts, double *d_jac, void *stream) {
  if (net < 0 || net >= NET_COUNT || !kCompiled[net]) {
    set_err("unsupported network id");
    return -1;
  }
  if (npts <= 0) return 0;
  if (!d_abundance || !d_rate || !d_density || !d_jac) {
    set_err("null device pointer");
    return -1;
  }
  const int neq = kNets[net].nspec - 1;
  hipStream_t s = (hipStream_t)stream;
  if (!hip_ok(hipMemsetAsync(d_jac, 0, (size_t)npts * neq * (neq + 1) * sizeof(double), s), "hipMemsetAsync"))
    return -1;
  JacRequest q{d_abundance, d_rate, d_density, npts, kNets[net].nspec, kNets[net].nreac,
               precision == PDRCHEM_MIXED ? 1 : 0, d_jac};
  const int rc = net == NET_REDUCED ? launch_wave_jacobian(net, q, s) : launch_wg_jacobian(net, q, s);
  if (rc != 0) {
    set_err("Jacobian kernel launch failed");
    return -1;
  }
  return 0;
}

void pdrchem_set_max_blocks(int max_blocks) { g_max_blocks = max_blocks < 0 ? 0 : max_blocks; }

void pdrchem_set_precision(int precision) { g_precision = precision == PDRCHEM_FP64 ? PDRCHEM_FP64 : PDRCHEM_MIXED; }

int pdrchem_network_id(int nspec, int nreac) {
  for (int i = 0; i < NET_COUNT; i++)
    if (kCompiled[i] && kNets[i].nspec == nspec && kNets[i].nreac == nreac) return i;
  return -1;
}

int pdrchem_kernel_info(int net, int precision, int *num_regs, int *lds_bytes, int *scratch_bytes,
                        int *blocks_per_cu) {
  KernelInfo ki{};
  if (net < 0 || net >= NET_COUNT || solver_info(net, precision == PDRCHEM_MIXED, &ki) != 0) {
    set_err("no solver kernel for this network");
    return -1;
  }
  if (num_regs) *num_regs = ki.num_regs;
  if (lds_bytes) *lds_bytes = ki.shared_bytes;
  if (scratch_bytes) *scratch_bytes = ki.local_bytes;
  if (blocks_per_cu) *blocks_per_cu = ki.blocks_per_cu;
  return 0;
}

int pdrchem_solve_device(int net, double *d_abundance, const double *d_rate, const double *d_density,
                         int npts, double rtol, double atol, double t_start, double t_end, int nelect,
                         int *d_status, int *d_stats, unsigned int *d_counter, int precision, void *stream) {
  if (net < 0 || net >= NET_COUNT || !kCompiled[net]) {
    set_err("unsupported network id");
    return -1;
  }
  if (npts <= 0) return 0;
  if (!d_abundance || !d_rate || !d_density || !d_status || !d_counter) {
    set_err("null device pointer");
    return -1;
  }
  if (nelect < 1 || nelect > kNets[net].nspec) {
    set_err("electron index out of range");
    return -1;
  }
  SolveRequest r{};
  r.abundance = d_abundance;
  r.rate = d_rate;
  r.density = d_density;
  r.npts = npts;
  r.nspec = kNets[net].nspec;
  r.nreac = kNets[net].nreac;
  r.rtol = rtol;
  r.atol = atol;
  r.t0 = t_start * SECONDS_IN_YEAR;
  r.tend = t_end * SECONDS_IN_YEAR;
  r.tout0 = 1.0e-4 * SECONDS_IN_YEAR;  // calculate_abundances.c:214
  r.nelect0 = nelect - 1;
  const SolverOptions opt = solver_options();
  r.mxstep = opt.mxstep;               // calculate_abundances.c:111 (1e7 unless set)
  r.xe_policy = opt.xe_policy;
  r.status = d_status;
  r.stats = d_stats;
  r.counter = d_counter;
  r.max_blocks = g_max_blocks;
  r.mixed = (precision == PDRCHEM_MIXED) ? 1 : 0;
  int rc = launch_solver(net, r, (hipStream_t)stream);
  if (rc != 0) {
    set_err("solver launch failed");
    return -1;
  }
  return 0;
}

int pdrchem_rates_device(int net, const pdrchem_phys *phys, const double *d_T, const double *d_Td, int nrays,
                         const double *d_G, const double *d_Av, const double *d_cols, int npts, double *d_rate,
                         int *d_neg, void *stream) {
  return pdrchem_rates_device_zeta(net, phys, d_T, d_Td, nrays, d_G, d_Av, d_cols, nullptr, npts, d_rate, d_neg,
                                   stream);
}

int pdrchem_rates_device_zeta(int net, const pdrchem_phys *phys, const double *d_T, const double *d_Td, int nrays,
                              const double *d_G, const double *d_Av, const double *d_cols, const double *d_zeta,
                              int npts, double *d_rate, int *d_neg, void *stream) {
  if (net < 0 || net >= NET_COUNT || !kCompiled[net]) {
    set_err("unsupported network id");
    return -1;
  }
  if (npts <= 0) return 0;
  if (!phys || !d_T || !d_Td || !d_G || !d_Av || !d_rate || !d_neg || nrays < 0) {
    set_err("bad argument to pdrchem_rates_device_zeta");
    return -1;
  }
  RateRequest q{};
  q.T = d_T;
  q.Td = d_Td;
  q.G = d_G;
  q.Av = d_Av;
  q.cols = d_cols;
  q.zeta_pts = d_zeta;
  q.npts = npts;
  q.nrays = nrays;
  q.ph = PhysParams{phys->zeta, phys->v_turb, phys->omega, phys->grain_radius, phys->metallicity, phys->g2d};
  q.rate = d_rate;
  q.neg = d_neg;
  if (launch_rates(net, q, (hipStream_t)stream) != 0) {
    set_err("rate kernel launch failed");
    return -1;
  }
  return 0;
}

int calculate_abundances_(double *abundance, double *rate, double *density, double *temperature, int *npart,
                          int *nspec, int *nreac) {
  (void)temperature;  // stored as T_g but unused by the RHS (odes_*.c)
  const double reltol = chemistry_module_mp_relative_abundance_tolerance_;
  const double abstol = chemistry_module_mp_absolute_abundance_tolerance_;
  const double start_time = maincode_module_mp_start_time_;
  const double end_time = maincode_module_mp_end_time_;
  int nelect = global_module_mp_nelect_;
  const int np = *npart, ns = *nspec, nr = *nreac;
  if (np <= 0) return 0;
  int net = pdrchem_network_id(ns, nr);
  if (net < 0) {
    set_err("no compiled network for (nspec, nreac)");
    return 1;
  }
  if (nelect <= 0) nelect = ns;  // electron is the last species (read_species.F90:164)
  std::lock_guard<std::mutex> lock(g_buf.mu);
  const size_t n_ab = (size_t)np * ns, n_rate = (size_t)np * nr;
  if (!g_buf.reserve(n_ab, n_rate, (size_t)np)) return 1;
  hipStream_t s = g_buf.stream;
  if (!hip_ok(hipMemcpyAsync(g_buf.ab, abundance, n_ab * sizeof(double), hipMemcpyHostToDevice, s), "H2D"))
    return 1;
  if (!hip_ok(hipMemcpyAsync(g_buf.rate, rate, n_rate * sizeof(double), hipMemcpyHostToDevice, s), "H2D"))
    return 1;
  if (!hip_ok(hipMemcpyAsync(g_buf.dens, density, (size_t)np * sizeof(double), hipMemcpyHostToDevice, s), "H2D"))
    return 1;
  if (!hip_ok(hipMemsetAsync(g_buf.counter, 0, sizeof(unsigned int), s), "memset")) return 1;
  if (pdrchem_solve_device(net, g_buf.ab, g_buf.rate, g_buf.dens, np, reltol, abstol, start_time, end_time,
                           nelect, g_buf.status, nullptr, g_buf.counter, host_precision(), s) != 0)
    return 1;
  std::vector<int> st((size_t)np);
  if (!hip_ok(hipMemcpyAsync(abundance, g_buf.ab, n_ab * sizeof(double), hipMemcpyDeviceToHost, s), "D2H"))
    return 1;
  if (!hip_ok(hipMemcpyAsync(st.data(), g_buf.status, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, s), "D2H"))
    return 1;
  if (!hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize")) return 1;
  int status = 0;
  for (int n = 0; n < np; n++)
    if (st[n] != 0) {
      status = st[n];
      break;
    }
  if (status != 0) log_failures(st, start_time, end_time);
  return status;
}

}  // extern "C"
