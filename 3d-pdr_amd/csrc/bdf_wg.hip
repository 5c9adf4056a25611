// Workgroup-per-system BDF kernels (networks of more than 32 equations),
// instantiated per network, and their launchers.  See bdf_wg_impl.h for the
// algorithm and layout.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include "wave_ops.h"
#include "gen/wg_full.h"
#include "gen/wg_mynetwork.h"
#include "pdr_internal.h"

namespace bdfg_full {
namespace NET = pdrwg_full;
#include "bdf_wg_impl.h"
}  // namespace bdfg_full

namespace bdfg_mynetwork {
namespace NET = pdrwg_mynetwork;
#include "bdf_wg_impl.h"
}  // namespace bdfg_mynetwork

namespace pdrchem {

namespace {

// Solver workspace (saved Jacobians, per-wavefront rate constants): one
// buffer per stream, grown on demand and kept for the life of the process.
// Launches on one stream run in order and may reuse it; launches on
// different streams may overlap and must not share it.
struct StreamWorkspace {
  struct Buf {
    void *ptr = nullptr;
    size_t bytes = 0;
  };
  std::mutex mu;
  std::map<hipStream_t, Buf> bufs;
  void *get(hipStream_t s, size_t need) {
    std::lock_guard<std::mutex> lock(mu);
    Buf &b = bufs[s];
    if (need > b.bytes) {
      // the previous buffer may still be in use by work queued on s
      if (b.ptr && (hipStreamSynchronize(s) != hipSuccess || hipFree(b.ptr) != hipSuccess)) return nullptr;
      b.ptr = nullptr;
      b.bytes = 0;
      if (hipMalloc(&b.ptr, need) != hipSuccess) return nullptr;
      b.bytes = need;
    }
    return b.ptr;
  }
};
StreamWorkspace g_ws;

}  // namespace

void *stream_workspace(hipStream_t s, size_t bytes) { return g_ws.get(s, bytes); }

namespace {

template <typename Args>
void fill_args(Args &a, const SolveRequest &r) {
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

template <typename Args>
int launch(int net, void (*kern)(Args, unsigned char *), const SolveRequest &r, size_t slot_bytes, hipStream_t s) {
  Args a;
  fill_args(a, r);
  int dev = 0, ncu = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0) != hipSuccess) return -1;
  if (per_cu < 1) per_cu = 1;
  LaunchConfig cfg = r.cfg;
  if (cfg.systems_per_block == 0 && cfg.blocks_per_cu == 0) cfg = tuned_config(net, r.mixed, r.npts);
  if (cfg.blocks_per_cu > 0 && per_cu > cfg.blocks_per_cu) per_cu = cfg.blocks_per_cu;
  long long full = (long long)ncu * per_cu;
  long long nb = r.npts < full ? r.npts : full;
  if (r.max_blocks > 0 && nb > r.max_blocks) nb = r.max_blocks;
  if (r.workspace && nb > (long long)(r.workspace_bytes / slot_bytes)) nb = (long long)(r.workspace_bytes / slot_bytes);
  if (nb < 1) {
    if (r.workspace) return -1;  // not even one workgroup's slot fits the caller's buffer
    nb = 1;
  }
  auto *ws = (unsigned char *)(r.workspace ? r.workspace : stream_workspace(s, (size_t)nb * slot_bytes));
  if (!ws) return -1;
  hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(256), 0, s, a, ws);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename Args>
size_t full_workspace(void (*kern)(Args, unsigned char *), size_t slot_bytes) {
  int dev = 0, ncu = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0) != hipSuccess) return 0;
  if (per_cu < 1) per_cu = 1;
  return (size_t)ncu * per_cu * slot_bytes;
}

template <typename Kern>
int info_of(Kern fn, KernelInfo *info) {
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

template <typename MT>
int launch_jac(void (*kern)(const double *, const double *, const double *, int, int, double *, unsigned char *),
               size_t slot_bytes, const JacRequest &q, hipStream_t s) {
  auto *ws = (unsigned char *)stream_workspace(s, (size_t)q.npts * slot_bytes);
  if (!ws) return -1;
  hipLaunchKernelGGL(kern, dim3((unsigned)q.npts), dim3(256), 0, s, q.x, q.rate, q.density, q.nspec, q.nreac, q.out,
                     ws);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

int launch_wg_jacobian(int net, const JacRequest &q, hipStream_t s) {
  switch (net) {
    case NET_FULL:
      return q.mixed ? launch_jac<float>(bdfg_full::wg_jac_kernel<float>, bdfg_full::Slot<float>::bytes, q, s)
                     : launch_jac<double>(bdfg_full::wg_jac_kernel<double>, bdfg_full::Slot<double>::bytes, q, s);
    case NET_MYNETWORK:
      return q.mixed
                 ? launch_jac<float>(bdfg_mynetwork::wg_jac_kernel<float>, bdfg_mynetwork::Slot<float>::bytes, q, s)
                 : launch_jac<double>(bdfg_mynetwork::wg_jac_kernel<double>, bdfg_mynetwork::Slot<double>::bytes, q,
                                      s);
    default:
      return -2;
  }
}

int launch_wg_solver(int net, const SolveRequest &r, hipStream_t s) {
  switch (net) {
    case NET_FULL:
      return r.mixed ? launch(net, bdfg_full::bdf_wg_kernel<float>, r, bdfg_full::Slot<float>::bytes, s)
                     : launch(net, bdfg_full::bdf_wg_kernel<double>, r, bdfg_full::Slot<double>::bytes, s);
    case NET_MYNETWORK:
      return r.mixed ? launch(net, bdfg_mynetwork::bdf_wg_kernel<float>, r, bdfg_mynetwork::Slot<float>::bytes, s)
                     : launch(net, bdfg_mynetwork::bdf_wg_kernel<double>, r, bdfg_mynetwork::Slot<double>::bytes, s);
    default:
      return -2;
  }
}

size_t wg_solver_workspace_bytes(int net, int mixed) {
  switch (net) {
    case NET_FULL:
      return mixed ? full_workspace(bdfg_full::bdf_wg_kernel<float>, bdfg_full::Slot<float>::bytes)
                   : full_workspace(bdfg_full::bdf_wg_kernel<double>, bdfg_full::Slot<double>::bytes);
    case NET_MYNETWORK:
      return mixed ? full_workspace(bdfg_mynetwork::bdf_wg_kernel<float>, bdfg_mynetwork::Slot<float>::bytes)
                   : full_workspace(bdfg_mynetwork::bdf_wg_kernel<double>, bdfg_mynetwork::Slot<double>::bytes);
    default:
      return 0;
  }
}

int wg_solver_info(int net, int mixed, KernelInfo *info) {
  switch (net) {
    case NET_FULL:
      return mixed ? info_of(bdfg_full::bdf_wg_kernel<float>, info) : info_of(bdfg_full::bdf_wg_kernel<double>, info);
    case NET_MYNETWORK:
      return mixed ? info_of(bdfg_mynetwork::bdf_wg_kernel<float>, info)
                   : info_of(bdfg_mynetwork::bdf_wg_kernel<double>, info);
    default:
      return -2;
  }
}

}  // namespace pdrchem
