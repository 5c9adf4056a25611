// Internal interfaces between the C-ABI layer and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace pdrchem {

enum NetId { NET_REDUCED = 0, NET_FULL = 1, NET_MYNETWORK = 2, NET_COUNT = 3 };

// launch shape of the persistent solver (0 = the default / the device's limit)
struct LaunchConfig {
  int systems_per_block;  // wave kernel: wavefronts (systems) per block, 1, 2 or 4
  int blocks_per_cu;      // cap on resident blocks per CU
};

// tuned launch shapes per (network, precision, batch-size bucket), written
// by the autotuner (autotune.cpp); a launch with cfg = {0, 0} uses the
// tuned shape of its bucket, else the default
LaunchConfig tuned_config(int net, int mixed, int npts);
int size_bucket(int npts);

// one batched solve: every pointer is device memory
struct SolveRequest {
  double *abundance;      // [npts][nspec], point-major (calculate_abundances.c:155)
  const double *rate;     // [npts][nreac]
  const double *density;  // [npts]
  int npts, nspec, nreac;
  double rtol, atol;
  double t0, tend, tout0; // seconds
  const double *tend_pts; // [npts] per-point end time (seconds) or null
  int nelect0;            // 0-based electron slot
  long long mxstep;       // steps per CVode call before flag -1 (calculate_abundances.c:111)
  int xe_policy;          // electron slot on return: 0 = x_e of the last RHS evaluation, 1 = charge sum
  int *status;            // [npts]
  int *stats;             // [npts][8] or null
  unsigned int *counter;  // work-queue counter, zero on entry
  int max_blocks;         // 0 = fill the device
  int mixed;              // 0: FP64 iteration matrix; 1: FP32 matrix, FP64 Newton
  LaunchConfig cfg;       // {0, 0}: tuned or default launch shape
  // solver workspace supplied by the caller (device memory of
  // workspace_bytes); null: the stream's workspace (allocated on demand).
  // With a caller's workspace the launch keeps to the resident blocks whose
  // workspace fits in it and allocates nothing.
  void *workspace;
  size_t workspace_bytes;
};

// library-wide solver options (pdrchem_set_mxstep, pdrchem_set_electron_policy)
struct SolverOptions {
  long long mxstep;
  int xe_policy;
};
SolverOptions solver_options();
// text returned by pdrchem_last_error on this thread
void set_last_error(const std::string &msg);

struct KernelInfo {
  int num_regs, shared_bytes, local_bytes, max_threads, blocks_per_cu;
};

// device workspace of at least `bytes` for solver launches on stream s
// (bdf_wg.hip); nullptr on allocation failure
void *stream_workspace(hipStream_t s, size_t bytes);

// reduced network: one system per wavefront (bdf_wave.hip)
int launch_wave_solver(int net, const SolveRequest &r, hipStream_t s);
int wave_solver_info(int net, int mixed, KernelInfo *info);
// full network and MYNETWORK: one system per workgroup (bdf_wg.hip)
int launch_wg_solver(int net, const SolveRequest &r, hipStream_t s);
int wg_solver_info(int net, int mixed, KernelInfo *info);
// workspace a launch of the whole device needs (0 if the network has no
// kernel of that kind); sizing the stream's workspace to it beforehand makes
// every later launch on the stream allocation-free (capturable)
size_t wave_solver_workspace_bytes(int net, int mixed);
size_t wg_solver_workspace_bytes(int net, int mixed);
// the solver's Jacobian at given abundances (introspection; one point per
// block): out[npts][nspec-1][nspec] in the bordered layout of the kernels
struct JacRequest {
  const double *x, *rate, *density;  // [npts][nspec], [npts][nreac], [npts]
  int npts, nspec, nreac, mixed;
  double *out;
};
int launch_wave_jacobian(int net, const JacRequest &q, hipStream_t s);
int launch_wg_jacobian(int net, const JacRequest &q, hipStream_t s);
// one unsigned int set to zero by a kernel (capturable, unlike a memset node)
int launch_zero_u32(unsigned int *p, hipStream_t s);

// rate evaluation (calc_reac_rates.F90)
struct PhysParams {
  double zeta, v_turb, omega, grain_radius, metallicity, g2d;
};
struct RateRequest {
  const double *T, *Td;   // [npts]
  const double *G, *Av;   // [npts][nrays]
  const double *cols;     // [npts][nrays][5] (H2, HD, CO, C, S) or null
  const double *zeta_pts; // [npts] cosmic-ray ionisation rate per point (units of ph.zeta) or null
  int npts, nrays;
  PhysParams ph;
  double *rate;           // [npts][nreac]
  int *neg;               // count of points with a negative rate (device int)
};
int launch_rates(int net, const RateRequest &r, hipStream_t s);

}  // namespace pdrchem
