// Device rate-coefficient kernels (calc_reac_rates.F90), one per network.
#include <hip/hip_runtime.h>

#include "gen/net_reduced.h"
#include "gen/shield_tables.h"
#include "gen/wg_full.h"
#include "gen/wg_mynetwork.h"
#include "pdr_internal.h"

namespace rates_reduced {
namespace NET = pdrnet_reduced;
#include "rates_impl.h"
}  // namespace rates_reduced

namespace rates_full {
namespace NET = pdrwg_full;
#include "rates_impl.h"
}  // namespace rates_full

namespace rates_mynetwork {
namespace NET = pdrwg_mynetwork;
#include "rates_impl.h"
}  // namespace rates_mynetwork

namespace pdrchem {

__global__ void zero_u32_kernel(unsigned int *p) { *p = 0u; }

int launch_zero_u32(unsigned int *p, hipStream_t s) {
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_rates(int net, const RateRequest &r, hipStream_t s) {
  long long total = 0;
  void (*kern)(RateRequest) = nullptr;
  switch (net) {
    case NET_REDUCED:
      total = (long long)r.npts * pdrnet_reduced::NREAC;
      kern = rates_reduced::rates_kernel;
      break;
    case NET_FULL:
      total = (long long)r.npts * pdrwg_full::NREAC;
      kern = rates_full::rates_kernel;
      break;
    case NET_MYNETWORK:
      total = (long long)r.npts * pdrwg_mynetwork::NREAC;
      kern = rates_mynetwork::rates_kernel;
      break;
    default:
      return -2;
  }
  long long nb = (total + 255) / 256;
  if (nb > 65536) nb = 65536;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(256), 0, s, r);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace pdrchem
