// Half-wavefront BDF kernels (two systems per wavefront), instantiated per
// network, and their launcher.  See bdf_pair_impl.h for the algorithm and
// layout.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "wave_ops.h"
#include "gen/net_reduced.h"
#include "pdr_internal.h"

namespace bdfp_reduced {
namespace NET = pdrnet_reduced;
#include "bdf_pair_impl.h"
}  // namespace bdfp_reduced

namespace pdrchem {

namespace {

template <typename MT, int NW>
int launch_pairs(const SolveRequest &r, hipStream_t s, int cap_per_cu) {
  bdfp_reduced::PointArgs a{};
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
  auto kern = bdfp_reduced::bdf_pair_kernel<MT, NW>;
  int dev = 0, ncu = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * NW, 0) != hipSuccess) return -1;
  if (per_cu < 1) per_cu = 1;
  if (cap_per_cu > 0 && per_cu > cap_per_cu) per_cu = cap_per_cu;
  const long long want = ((long long)r.npts + 2 * NW - 1) / (2 * NW);
  const long long full = (long long)ncu * per_cu;
  long long nb = want < full ? want : full;
  if (r.max_blocks > 0 && nb > r.max_blocks) nb = r.max_blocks;
  constexpr size_t per_block = 2 * NW * sizeof(bdfp_reduced::PairWS<MT>);
  if (r.workspace && nb > (long long)(r.workspace_bytes / per_block)) nb = (long long)(r.workspace_bytes / per_block);
  if (nb < 1) {
    if (r.workspace) return -1;
    nb = 1;
  }
  auto *ws = (bdfp_reduced::PairWS<MT> *)(r.workspace ? r.workspace : stream_workspace(s, (size_t)nb * per_block));
  if (!ws) return -1;
  hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(64 * NW), 0, s, a, ws);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

int launch_pair_solver(int net, const SolveRequest &r, int nw, hipStream_t s) {
  if (net != NET_REDUCED || r.published) return -2;
  const int cap = r.cfg.blocks_per_cu;
  if (r.mixed) return nw == 2 ? launch_pairs<float, 2>(r, s, cap) : launch_pairs<float, 4>(r, s, cap);
  return nw == 2 ? launch_pairs<double, 2>(r, s, cap) : launch_pairs<double, 4>(r, s, cap);
}

size_t pair_solver_workspace_bytes(int net, int mixed) {
  if (net != NET_REDUCED) return 0;
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  auto one = [&](auto kern, int nw, size_t sys_bytes) -> size_t {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * nw, 0) != hipSuccess) return 0;
    return (size_t)ncu * std::max(per_cu, 1) * 2 * nw * sys_bytes;
  };
  if (mixed)
    return std::max(one(bdfp_reduced::bdf_pair_kernel<float, 2>, 2, sizeof(bdfp_reduced::PairWS<float>)),
                    one(bdfp_reduced::bdf_pair_kernel<float, 4>, 4, sizeof(bdfp_reduced::PairWS<float>)));
  return std::max(one(bdfp_reduced::bdf_pair_kernel<double, 2>, 2, sizeof(bdfp_reduced::PairWS<double>)),
                  one(bdfp_reduced::bdf_pair_kernel<double, 4>, 4, sizeof(bdfp_reduced::PairWS<double>)));
}

}  // namespace pdrchem
