// Wavefront-per-system BDF kernels, instantiated per network, and their
// launchers.  See bdf_wave_impl.h for the algorithm and layout.
#include <hip/hip_runtime.h>

#include "wave_ops.h"
#include "gen/net_reduced.h"
#include "pdr_internal.h"

namespace bdfw_reduced {
namespace NET = pdrnet_reduced;
#include "bdf_wave_impl.h"
}  // namespace bdfw_reduced

namespace pdrchem {

template <typename Args>
static void fill_args(Args &a, const SolveRequest &r) {
  a.abundance = r.abundance;
  a.rate = r.rate;
  a.density = r.density;
  a.nsys = r.npts;
  a.ab_stride = r.nspec;
  a.rate_stride = r.nreac;
  a.rtol = r.rtol;
  a.atol = r.atol;
  a.t0 = r.t0;
  a.tend = r.tend;
  a.tend_pts = r.tend_pts;
  a.tout0 = r.tout0;
  a.nelect0 = r.nelect0;
  a.mxstep = r.mxstep;
  a.status = r.status;
  a.stats = r.stats;
  a.counter = r.counter;
}

template <typename MT>
static int launch_persistent(void (*kern)(bdfw_reduced::PointArgs, bdfw_reduced::WaveWS<MT> *),
                             const bdfw_reduced::PointArgs &a, int npts, hipStream_t s, int max_blocks) {
  constexpr int WAVES = bdfw_reduced::WAVES;
  int dev = 0, ncu = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * WAVES, 0) != hipSuccess) return -1;
  if (per_cu < 1) per_cu = 1;
  long long want = ((long long)npts + WAVES - 1) / WAVES;
  long long full = (long long)ncu * per_cu;
  long long nb = want < full ? want : full;
  if (max_blocks > 0 && nb > max_blocks) nb = max_blocks;
  if (nb < 1) nb = 1;
  auto *ws = (bdfw_reduced::WaveWS<MT> *)stream_workspace(s, (size_t)nb * WAVES * sizeof(bdfw_reduced::WaveWS<MT>));
  if (!ws) return -1;
  hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(64 * WAVES), 0, s, a, ws);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_wave_solver(int net, const SolveRequest &r, hipStream_t s) {
  switch (net) {
    case NET_REDUCED: {
      bdfw_reduced::PointArgs a;
      fill_args(a, r);
      if (r.mixed) return launch_persistent<float>(bdfw_reduced::bdf_wave_kernel<float>, a, r.npts, s, r.max_blocks);
      return launch_persistent<double>(bdfw_reduced::bdf_wave_kernel<double>, a, r.npts, s, r.max_blocks);
    }
    default:
      return -2;
  }
}

template <typename MT>
static size_t full_workspace(void (*kern)(bdfw_reduced::PointArgs, bdfw_reduced::WaveWS<MT> *)) {
  constexpr int WAVES = bdfw_reduced::WAVES;
  int dev = 0, ncu = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * WAVES, 0) != hipSuccess) return 0;
  if (per_cu < 1) per_cu = 1;
  return (size_t)ncu * per_cu * WAVES * sizeof(bdfw_reduced::WaveWS<MT>);
}

size_t wave_solver_workspace_bytes(int net, int mixed) {
  if (net != NET_REDUCED) return 0;
  return mixed ? full_workspace<float>(bdfw_reduced::bdf_wave_kernel<float>)
               : full_workspace<double>(bdfw_reduced::bdf_wave_kernel<double>);
}

template <typename Kern>
static int kernel_info(Kern fn, KernelInfo *info) {
  hipFuncAttributes attr;
  if (hipFuncGetAttributes(&attr, (const void *)fn) != hipSuccess) return -1;
  info->num_regs = attr.numRegs;
  info->shared_bytes = (int)attr.sharedSizeBytes;
  info->local_bytes = (int)attr.localSizeBytes;
  info->max_threads = attr.maxThreadsPerBlock;
  int per_cu = 0;
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0);
  info->blocks_per_cu = per_cu;
  return 0;
}

int wave_solver_info(int net, int mixed, KernelInfo *info) {
  switch (net) {
    case NET_REDUCED:
      return mixed ? kernel_info(bdfw_reduced::bdf_wave_kernel<float>, info)
                   : kernel_info(bdfw_reduced::bdf_wave_kernel<double>, info);
    default:
      return -2;
  }
}

}  // namespace pdrchem
