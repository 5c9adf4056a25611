// Allocation-free variant of pdrchem_solve_device (no counterpart in the
// reference): for callers that cannot allocate device memory -- code under
// HIP graph capture, tight serving loops, hosts that forbid hipMalloc.
//
// The solver's scratch memory -- the work-queue counter and the per-block
// workspace of saved Jacobians and slot rate constants, which
// pdrchem_solve_device takes from a per-stream buffer allocated on first use
// -- lives here in one file-scope device array of fixed size, present on
// every GPU the library runs on until the process ends.  Its device address
// is looked up once per device (hipGetSymbolAddress, on the first call with
// that device current) and only that address is handed to kernels.
//
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// (two streams, two host threads) race on it and must be serialised by the
// caller.  The counter is cleared by a kernel at the start of each call, so
// the call is capturable; nothing else needs to start at zero (the workspace
// is written before it is read within each point's solve).
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/pdrchem.h"
#include "gen/net_reduced.h"
#include "gen/wg_full.h"
#include "gen/wg_mynetwork.h"
#include "pdr_internal.h"

using namespace pdrchem;

namespace {

// 48 MiB: the counter (first 256 bytes), then the solver workspace.  A
// whole-device launch on MI355X (256 CUs) needs at most 34.6 MB (reduced
// network, FP64 form: 768 blocks of four 11,264-byte wavefront workspaces);
// the full network needs 46.1 MB (FP32 matrix, 512 workgroups of 90,112
// bytes) or 39.8 MB (FP64, 256 of 155,648).  A device on which a whole-device
// launch would need more runs the persistent kernel on as many blocks as fit.
constexpr size_t kScratchBytes = size_t(48) << 20;
constexpr size_t kCounterBytes = 256;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

// the array's device address on the current device, looked up once
void *scratch_ptr(int dev) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return nullptr;
    }
    g_ptr[dev] = p;
  }
  return g_ptr[dev];
}

const int kNspec[NET_COUNT] = {pdrnet_reduced::NSPEC, pdrwg_full::NSPEC, pdrwg_mynetwork::NSPEC};
const int kNreac[NET_COUNT] = {pdrnet_reduced::NREAC, pdrwg_full::NREAC, pdrwg_mynetwork::NREAC};

}  // namespace

extern "C" {

int pdrchem_solve_device_noalloc(int net, double *d_abundance, const double *d_rate, const double *d_density,
                                 int npts, double rtol, double atol, double t_start, double t_end, int nelect,
                                 int *d_status, int *d_stats, int precision, void *stream) {
  if (net < 0 || net >= NET_COUNT) {
    set_last_error("unsupported network id");
    return -1;
  }
  if (npts < 0) {
    set_last_error("negative point count");
    return -1;
  }
  if (npts == 0) return 0;
  if (!d_abundance || !d_rate || !d_density || !d_status) {
    set_last_error("null device pointer");
    return -1;
  }
  if (nelect < 1 || nelect > kNspec[net]) {
    set_last_error("electron index out of range");
    return -1;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_last_error("no current device (or device index beyond the scratch table)");
    return -1;
  }
  unsigned char *p = (unsigned char *)scratch_ptr(dev);
  if (!p) return -1;
  hipStream_t s = (hipStream_t)stream;
  SolveRequest r{};
  r.abundance = d_abundance;
  r.rate = d_rate;
  r.density = d_density;
  r.npts = npts;
  r.nspec = kNspec[net];
  r.nreac = kNreac[net];
  r.rtol = rtol;
  r.atol = atol;
  r.t0 = t_start * 3.1556926e7;     // calculate_abundances.c:90
  r.tend = t_end * 3.1556926e7;
  r.tout0 = 1.0e-4 * 3.1556926e7;   // calculate_abundances.c:214
  r.nelect0 = nelect - 1;
  const SolverOptions opt = solver_options();
  r.mxstep = opt.mxstep;
  r.xe_policy = opt.xe_policy;
  r.status = d_status;
  r.stats = d_stats;
  r.counter = (unsigned int *)p;
  r.max_blocks = 0;
  r.mixed = precision == PDRCHEM_MIXED ? 1 : 0;
  r.workspace = p + kCounterBytes;
  r.workspace_bytes = kScratchBytes - kCounterBytes;
  if (launch_zero_u32(r.counter, s) != 0) {
    set_last_error("work-queue reset failed");
    return -1;
  }
  const int rc = net == NET_REDUCED ? launch_wave_solver(net, r, s) : launch_wg_solver(net, r, s);
  if (rc != 0) {
    set_last_error("solver launch failed (or its workspace does not fit the 48 MiB scratch)");
    return -1;
  }
  return 0;
}

}  // extern "C"
