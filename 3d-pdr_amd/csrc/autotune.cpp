// Autotuner of the persistent solver's launch shape (pdrchem_autotune).
//
// The choices: for the wavefront-per-system kernel (reduced network), how
// many systems (wavefronts) share a block -- 1, 2 or 4; a block's wavefronts
// share one copy of the RHS tables in LDS, so 4 packs the most systems per
// CU, while 1 spreads a small batch over more CUs -- and a cap on resident
// blocks per CU; for the workgroup-per-system kernel (full network,
// MYNETWORK), 1 or 2 resident workgroups per CU.  The autotuner runs every
// candidate on the caller's own batch (device data, restored between runs
// and at the end), times each with HIP events, and records the fastest for
// the batch's size bucket (powers of 4 of the point count) and precision.
// Later launches of that network and precision -- pdrchem_solve_device,
// calculate_abundances_, the request loop, the grid cloud -- use the shape of
// their bucket, or of the largest tuned bucket below theirs.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "../../include/pdrchem.h"
#include "gen/net_reduced.h"
#include "gen/wg_full.h"
#include "gen/wg_mynetwork.h"
#include "pdr_internal.h"

namespace pdrchem {
namespace {

constexpr int kBuckets = 16;
struct Entry {
  bool set = false;
  LaunchConfig cfg{0, 0};
  double ms = 0.0;
};
std::mutex g_mu;
Entry g_tab[NET_COUNT][2][kBuckets];

const int kNspec[NET_COUNT] = {pdrnet_reduced::NSPEC, pdrwg_full::NSPEC, pdrwg_mynetwork::NSPEC};
const int kNreac[NET_COUNT] = {pdrnet_reduced::NREAC, pdrwg_full::NREAC, pdrwg_mynetwork::NREAC};

}  // namespace

// bucket b holds point counts in (4^(b-1), 4^b]
int size_bucket(int npts) {
  int b = 0;
  long long v = 1;
  while (v < npts && b < kBuckets - 1) {
    v *= 4;
    b++;
  }
  return b;
}

// The bucket's own shape if tuned; else that of the largest tuned bucket
// below it (beyond a few thousand points the device is saturated and the
// ranking of shapes no longer changes with the batch size); else the default.
LaunchConfig tuned_config(int net, int mixed, int npts) {
  if (net < 0 || net >= NET_COUNT) return LaunchConfig{0, 0};
  std::lock_guard<std::mutex> lock(g_mu);
  for (int b = size_bucket(npts); b >= 0; b--) {
    const Entry &e = g_tab[net][mixed ? 1 : 0][b];
    if (e.set) return e.cfg;
  }
  return LaunchConfig{0, 0};
}

}  // namespace pdrchem

using namespace pdrchem;

extern "C" {

int pdrchem_autotune(int net, int precision, double *d_abundance, const double *d_rate, const double *d_density,
                     int npts, double t_end, int nelect, void *stream, int *systems_per_block, int *blocks_per_cu,
                     double *best_ms) {
  if (net < 0 || net >= NET_COUNT || npts <= 0 || !d_abundance || !d_rate || !d_density || nelect < 1 ||
      nelect > kNspec[net] || !(t_end > 0) || (precision != PDRCHEM_FP64 && precision != PDRCHEM_MIXED))
    return -1;
  hipStream_t s = (hipStream_t)stream;
  const int mixed = precision == PDRCHEM_MIXED ? 1 : 0;
  std::vector<LaunchConfig> cand;
  if (net == NET_REDUCED)
    cand = {{4, 0}, {2, 0}, {1, 0}, {4, 2}};
  else
    cand = {{1, 0}, {1, 1}};
  const size_t abytes = (size_t)npts * kNspec[net] * sizeof(double);
  double *backup = nullptr;
  int *status = nullptr;
  unsigned int *counter = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = -1;
  double best = -1.0;
  LaunchConfig best_cfg{0, 0};
  bool backed_up = false;
  do {
    if (hipMalloc(&backup, abytes) != hipSuccess || hipMalloc(&status, (size_t)npts * sizeof(int)) != hipSuccess ||
        hipMalloc(&counter, sizeof(unsigned int)) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
      break;
    if (hipMemcpyAsync(backup, d_abundance, abytes, hipMemcpyDeviceToDevice, s) != hipSuccess) break;
    backed_up = true;
    SolveRequest r{};
    r.abundance = d_abundance;
    r.rate = d_rate;
    r.density = d_density;
    r.nspec = kNspec[net];
    r.nreac = kNreac[net];
    r.rtol = chemistry_module_mp_relative_abundance_tolerance_;
    r.atol = chemistry_module_mp_absolute_abundance_tolerance_;
    r.t0 = 0.0;
    r.tend = t_end * 3.1556926e7;  // calculate_abundances.c:90
    r.tout0 = 1.0e-4 * 3.1556926e7;
    r.nelect0 = nelect - 1;
    r.mxstep = solver_options().mxstep;  // calculate_abundances.c:111 (1e7 unless set)
    r.xe_policy = solver_options().xe_policy;
    r.status = status;
    r.counter = counter;
    r.mixed = mixed;
    bool ok = true;
    for (const LaunchConfig &c : cand) {
      r.cfg = c;
      // warm-up on a few points (first launch of a shape), then the timed run
      for (int pass = 0; pass < 2 && ok; pass++) {
        r.npts = pass == 0 ? (npts < 64 ? npts : 64) : npts;
        ok = hipMemcpyAsync(d_abundance, backup, abytes, hipMemcpyDeviceToDevice, s) == hipSuccess &&
             hipMemsetAsync(counter, 0, sizeof(unsigned int), s) == hipSuccess &&
             hipEventRecord(e0, s) == hipSuccess &&
             (net == NET_REDUCED ? launch_wave_solver(net, r, s) : launch_wg_solver(net, r, s)) == 0 &&
             hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
      }
      if (!ok) break;
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        ok = false;
        break;
      }
      if (best < 0 || ms < best) {
        best = ms;
        best_cfg = c;
      }
    }
    if (!ok) break;
    if (hipMemcpyAsync(d_abundance, backup, abytes, hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      break;
    {
      std::lock_guard<std::mutex> lock(g_mu);
      Entry &e = g_tab[net][mixed][size_bucket(npts)];
      e.set = true;
      e.cfg = best_cfg;
      e.ms = best;
    }
    rc = 0;
  } while (false);
  // on a failure after the backup was taken, the caller's abundances may be
  // partly integrated: put them back before the backup goes away
  if (rc != 0 && backed_up) {
    (void)hipMemcpyAsync(d_abundance, backup, abytes, hipMemcpyDeviceToDevice, s);
    (void)hipStreamSynchronize(s);
  } else if (rc != 0) {
    (void)hipStreamSynchronize(s);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (counter) (void)hipFree(counter);
  if (status) (void)hipFree(status);
  if (backup) (void)hipFree(backup);
  if (rc == 0) {
    if (systems_per_block) *systems_per_block = best_cfg.systems_per_block;
    if (blocks_per_cu) *blocks_per_cu = best_cfg.blocks_per_cu;
    if (best_ms) *best_ms = best;
  }
  return rc;
}

void pdrchem_clear_tuning(void) {
  std::lock_guard<std::mutex> lock(g_mu);
  for (auto &a : g_tab)
    for (auto &b : a)
      for (Entry &e : b) e = Entry{};
}

int pdrchem_set_tuned_config(int net, int precision, int npts, int systems_per_block, int blocks_per_cu) {
  if (net < 0 || net >= NET_COUNT || npts <= 0 || systems_per_block < 0 || systems_per_block > 4 ||
      systems_per_block == 3 || blocks_per_cu < 0)
    return -1;
  std::lock_guard<std::mutex> lock(g_mu);
  Entry &e = g_tab[net][precision == PDRCHEM_MIXED ? 1 : 0][size_bucket(npts)];
  e.set = systems_per_block != 0 || blocks_per_cu != 0;
  e.cfg = LaunchConfig{systems_per_block, blocks_per_cu};
  e.ms = 0.0;
  return 0;
}

int pdrchem_tuned_config(int net, int precision, int npts, int *systems_per_block, int *blocks_per_cu) {
  if (net < 0 || net >= NET_COUNT || npts <= 0) return -1;
  const LaunchConfig c = tuned_config(net, precision == PDRCHEM_MIXED ? 1 : 0, npts);
  if (systems_per_block) *systems_per_block = c.systems_per_block;
  if (blocks_per_cu) *blocks_per_cu = c.blocks_per_cu;
  return (c.systems_per_block || c.blocks_per_cu) ? 1 : 0;  // 1: a tuned shape applies
}

}  // extern "C"
