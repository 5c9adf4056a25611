// Wavefront-per-system BDF kernels, instantiated per network, and their
// launchers.  See bdf_wave_impl.h for the algorithm and layout.
#include <hip/hip_runtime.h>

#include <algorithm>

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
  a.xe_policy = r.xe_policy;
  a.status = r.status;
  a.stats = r.stats;
  a.counter = r.counter;
  a.published = r.published;
  a.stop = r.stop;
  a.ring_mask = r.ring_mask;
  a.idle_ticks = r.idle_ticks;
  a.life_ticks = r.life_ticks;
}

template <typename MT, int NW>
static int launch_persistent(const bdfw_reduced::PointArgs &a, int npts, hipStream_t s, int max_blocks,
                             int cap_per_cu, void *wsp, size_t ws_bytes) {
  auto kern = bdfw_reduced::bdf_wave_kernel<MT, NW>;
  int dev = 0, ncu = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * NW, 0) != hipSuccess) return -1;
  if (per_cu < 1) per_cu = 1;
  if (cap_per_cu > 0 && per_cu > cap_per_cu) per_cu = cap_per_cu;
  long long want = ((long long)npts + NW - 1) / NW;
  long long full = (long long)ncu * per_cu;
  long long nb = want < full ? want : full;
  if (max_blocks > 0 && nb > max_blocks) nb = max_blocks;
  constexpr size_t per_block = NW * sizeof(bdfw_reduced::WaveWS<MT>);
  if (wsp && nb > (long long)(ws_bytes / per_block)) nb = (long long)(ws_bytes / per_block);
  if (nb < 1) {
    if (wsp) return -1;  // not even one block's workspace fits the caller's buffer
    nb = 1;
  }
  auto *ws = (bdfw_reduced::WaveWS<MT> *)(wsp ? wsp : stream_workspace(s, (size_t)nb * per_block));
  if (!ws) return -1;
  hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(64 * NW), 0, s, a, ws);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename MT>
static int launch_shape(const bdfw_reduced::PointArgs &a, int npts, hipStream_t s, int max_blocks,
                        LaunchConfig cfg, void *wsp, size_t ws_bytes) {
  switch (cfg.systems_per_block) {
    case 1: return launch_persistent<MT, 1>(a, npts, s, max_blocks, cfg.blocks_per_cu, wsp, ws_bytes);
    case 2: return launch_persistent<MT, 2>(a, npts, s, max_blocks, cfg.blocks_per_cu, wsp, ws_bytes);
    default: return launch_persistent<MT, 4>(a, npts, s, max_blocks, cfg.blocks_per_cu, wsp, ws_bytes);
  }
}

int launch_wave_solver(int net, const SolveRequest &r, hipStream_t s) {
  switch (net) {
    case NET_REDUCED: {
      bdfw_reduced::PointArgs a;
      fill_args(a, r);
      LaunchConfig cfg = r.cfg;
      if (cfg.systems_per_block == 0 && cfg.blocks_per_cu == 0) cfg = tuned_config(net, r.mixed, r.npts);
      if (r.mixed) return launch_shape<float>(a, r.npts, s, r.max_blocks, cfg, r.workspace, r.workspace_bytes);
      return launch_shape<double>(a, r.npts, s, r.max_blocks, cfg, r.workspace, r.workspace_bytes);
    }
    default:
      return -2;
  }
}

int launch_wave_jacobian(int net, const JacRequest &q, hipStream_t s) {
  if (net != NET_REDUCED) return -2;
  if (q.mixed) {
    auto *ws = (bdfw_reduced::WaveWS<float> *)stream_workspace(s, (size_t)q.npts * sizeof(bdfw_reduced::WaveWS<float>));
    if (!ws) return -1;
    hipLaunchKernelGGL(bdfw_reduced::wave_jac_kernel<float>, dim3((unsigned)q.npts), dim3(64), 0, s, q.x, q.rate,
                       q.density, q.nspec, q.nreac, q.out, ws);
  } else {
    auto *ws =
        (bdfw_reduced::WaveWS<double> *)stream_workspace(s, (size_t)q.npts * sizeof(bdfw_reduced::WaveWS<double>));
    if (!ws) return -1;
    hipLaunchKernelGGL(bdfw_reduced::wave_jac_kernel<double>, dim3((unsigned)q.npts), dim3(64), 0, s, q.x, q.rate,
                       q.density, q.nspec, q.nreac, q.out, ws);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// the largest whole-device workspace over the block shapes
template <typename MT, int NW>
static size_t shape_workspace(int ncu) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bdfw_reduced::bdf_wave_kernel<MT, NW>, 64 * NW, 0) !=
      hipSuccess)
    return 0;
  return (size_t)ncu * std::max(per_cu, 1) * NW * sizeof(bdfw_reduced::WaveWS<MT>);
}

template <typename MT>
static size_t full_workspace() {
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  return std::max(shape_workspace<MT, 1>(ncu), std::max(shape_workspace<MT, 2>(ncu), shape_workspace<MT, 4>(ncu)));
}

size_t wave_solver_workspace_bytes(int net, int mixed) {
  if (net != NET_REDUCED) return 0;
  return mixed ? full_workspace<float>() : full_workspace<double>();
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
      return mixed ? kernel_info(bdfw_reduced::bdf_wave_kernel<float, 4>, info)
                   : kernel_info(bdfw_reduced::bdf_wave_kernel<double, 4>, info);
    default:
      return -2;
  }
}

}  // namespace pdrchem
