// Internal interfaces between the C-ABI layer and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

struct pdrchem_coolant_tables;

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
  // standing queue (all null/zero for a closed batch): the point arrays are
  // a ring of ring_mask + 1 points in host memory; published and stop are
  // host words, idle/life bounds in real-time clock ticks (PointArgs)
  const unsigned int *published;
  const unsigned int *stop;
  unsigned int ring_mask;
  unsigned long long idle_ticks, life_ticks;
};

// status of a standing-queue point not yet solved (the host writes it before
// publishing the point; the solver's statuses are small integers)
constexpr int kPendingStatus = -2147483647 - 1;

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

// ---------------------------------------------------------------------------
// thermal balance (thermal.hip): coolant tables, level populations, the
// temperature update
// ---------------------------------------------------------------------------
// one coolant's tables on the device, in the reference's column-major layout
// (read_input.F90): X(i, j) at [i + j*nlev], TEMPERATURES(p, k) at
// [p + 7k], partner p's COL(i, j, k) at col[p*nlev*nlev*ntemp + i + j*nlev +
// k*nlev*nlev]; partners H2, pH2, oH2, e, H, He, H+
struct CoolantDev {
  int nlev, ntemp;
  int off;  // first level of this coolant in the per-point population rows
  const double *E, *g, *A, *B, *F, *temps, *col;
};
struct LvgTables {
  static constexpr int kMaxLev = 41;
  int ncool, nlevtot;
  CoolantDev c[4];   // CII, CI, OI, CO
  int species[4];    // 0-based abundance slots of C+, C, O, CO
  int partner[5];    // 0-based abundance slots of H, H+, e-, He, H2
};
struct LvgArgs {
  static constexpr int kLines = 17;  // the lines of the reference's .line.fin (3DPDR.F90:1407-1416)
  int npts, nspec, nrays, pseudo_ray;  // pseudo_ray >= 0: PSEUDO_1D, optical depth along that ray only
  const double *T, *Td, *rho, *ab;
  const double *pop;                   // [npts][nlevtot] current populations
  const double *S;                     // [npts][nrays][nlevtot] level columns
  const unsigned char *skip;           // points not solved (their outputs are left)
  double v_turb, metallicity, g2d;
  int co_fix, co_average_co, levpop_iteration;
  double *solution;                    // [npts][nlevtot]
  double *cool;                        // [npts][4]
  double *lines;                       // [npts][kLines]
  // optical depth of each line along ray tau_ray (NULL: not wanted); 0 for a
  // line without one (empty upper level, |tpop| < 1e-50)
  double *tau_lines;                   // [npts][kLines]
  int tau_ray;
  // index of (upper, lower) of coolant ci among the output lines, or -1:
  // CII 2-1; CI 2-1, 3-1, 3-2; OI 2-1, 3-1, 3-2; CO J+1 -> J for J = 0..9
  __host__ __device__ static int line_slot(int ci, int il, int jl) {
    if (ci == 0) return (il == 1 && jl == 0) ? 0 : -1;
    if (ci == 1 || ci == 2) {
      const int b = ci == 1 ? 1 : 4;
      if (il == 1 && jl == 0) return b;
      if (il == 2 && jl == 0) return b + 1;
      if (il == 2 && jl == 1) return b + 2;
      return -1;
    }
    return (il == jl + 1 && jl < 10) ? 7 + jl : -1;
  }
};
// per-point state of the temperature search (3DPDR.F90 arrays of the same
// names): previous temperature, bracket, last change ('H' heated, 'C'
// cooled, 0 none), flags
struct ThermalPoint {
  double Tprev, Tlow, Thigh;
  char prev;
  unsigned char binarychop, converged, doleveltmin;
};
struct ThermalArgs {
  int nact;
  const int *idx;              // compact index -> point
  const double *heat_c;        // [nact][12] heating terms of the compact points
  double *heat;                // [npts][12]
  const double *total_cool;    // [npts]
  double *fmean, *fratio;      // [npts]
  int level_conv, first_time, temp_fix;
  double Tmin, Tmax, Fcrit, Tdiff;
  ThermalPoint *state;         // [npts]
  double *T;                   // [npts] gas temperature
  unsigned char *level_converged;  // [npts]
};
int launch_dust_temperature(int npts, int nrays, const double *G, const double *uv, double *Td, hipStream_t s);
int launch_lte(const LvgTables &t, int npts, const double *T, const double *ab, int nspec, const double *rho,
               const unsigned char *skip, double *pop, hipStream_t s);
int launch_level_columns(int npts, int nrays, int ray0, int nray_sel, const long long *offs, const int *count,
                         const int *proj, const double *step, const double *pop, int nlevtot, double *S,
                         hipStream_t s);
int launch_lvg(const LvgTables &t, const LvgArgs &a, int coolant, hipStream_t s);
int launch_total_cool(int npts, const double *cool, double *total, hipStream_t s);
int launch_thermal_update(const ThermalArgs &a, hipStream_t s);
int launch_relch(const LvgTables &t, int npts, const double *ab, int nspec, const double *sol, const double *pop,
                 unsigned char *level_converged, unsigned char *cconv, unsigned *counts, hipStream_t s);
int launch_count_converged(int npts, const ThermalPoint *st, unsigned *count, hipStream_t s);
int launch_copy(long long n, const double *src, double *dst, hipStream_t s);
int launch_init_thermal_state(int npts, const double *T, double Tmin, double Tmax, ThermalPoint *st, hipStream_t s);
int launch_skip_flags(int npts, const ThermalPoint *st, const unsigned char *lc, int with_levels,
                      unsigned char *skip, hipStream_t s);
// the device tables behind a pdrchem_coolant_tables handle (thermal.hip)
const LvgTables *coolant_tables_dev(const ::pdrchem_coolant_tables *t);
// heating terms (heating.hip): see pdrchem_heating_device
int launch_heating(int net, const PhysParams &ph, const int *index, int npts, const double *n, const double *T,
                   const double *Td, const double *uv, const double *ab, const double *rate, const double *zeta,
                   double *heat, hipStream_t s);

}  // namespace pdrchem
