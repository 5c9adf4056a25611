/*
 * pdrchem.h -- C ABI of libpdrchem.so, the MI355X-native per-grid-point
 * chemistry solve for 3D-PDR.
 *
 * Drop-in entry point: calculate_abundances_ replaces the reference's
 * CVODE-based implementation with the same symbol, arguments and in-place
 * semantics; a Fortran host keeps its implicit-interface CALL.  Everything
 * else is additive: batched device-pointer entry points for callers that keep
 * state in HBM, device rate evaluation, and introspection.
 *
 * No function here takes or returns a torch type; pointers are plain C
 * pointers, sizes are int.
 */
#ifndef PDRCHEM_H
#define PDRCHEM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------
 * Reference interface
 * ------------------------------------------------------------------- */

/* Replaces calculate_abundances_ (calculate_abundances.c:64-65), called from
 * 3DPDR.F90:498, :623, :1929 as
 *   CALL CALCULATE_ABUNDANCES(ABUNDANCE, RATE, DENSITY, TEMPERATURE, NPART,
 *                             NSPEC, NREAC)
 * abundance[npart*nspec]  in/out, point-major (index n*nspec+i); on return
 *                         abundances < atol are 0 and slot nelect-1 holds
 *                         x_e from the last RHS evaluation
 *                         (calculate_abundances.c:308-317)
 * rate[npart*nreac]       rate coefficients per point (read only)
 * density[npart]          n_H per point (read only)
 * temperature[npart]      gas temperature (read only; unused by the RHS)
 * Tolerances, start/end time (years) and the electron index are read from
 * the globals below, as the reference does (calculate_abundances.c:78-91).
 * The network is chosen at run time from (nspec, nreac).
 * Returns 0 on success, 1 on a set-up failure (unsupported network, device
 * error), else the CVODE-style flag of a failed point (-1 too much work,
 * -3 error-test failures, -4 convergence failures, ...), like the reference. */
int calculate_abundances_(double *abundance, double *rate, double *density, double *temperature,
                          int *npart, int *nspec, int *nreac);

/* Globals read by calculate_abundances_.  In the reference they are owned by
 * Fortran modules with bind(c) names (modules.F90:165-166, 383, 529-530); the
 * library carries weak definitions with the reference's defaults
 * (3DPDR.F90:452, 462-463; params.dat end time 1e7 yr) that a Fortran host's
 * definitions replace at link/load time. */
extern double chemistry_module_mp_relative_abundance_tolerance_;
extern double chemistry_module_mp_absolute_abundance_tolerance_;
extern double maincode_module_mp_start_time_;
extern double maincode_module_mp_end_time_;
extern int global_module_mp_nelect_;

/* ---------------------------------------------------------------------
 * Batched device interface (no counterpart in the reference; used by hosts
 * that keep state resident in HBM and by the benchmark)
 * ------------------------------------------------------------------- */

/* network id for (nspec, nreac); -1 if no compiled kernel matches */
int pdrchem_network_id(int nspec, int nreac);

/* precision of the iteration matrix (the residual, Newton update and error
 * test are FP64 in both) */
#define PDRCHEM_FP64 0   /* plain form: J and its scaled inverse in FP64 */
#define PDRCHEM_MIXED 1  /* J and the scaled inverse in FP32 */

/* Solve npts independent points already in device memory, asynchronously
 * on `stream` (a hipStream_t; NULL = default stream).  Same per-point
 * semantics as calculate_abundances_.  t_start/t_end in years; nelect is
 * 1-based.  d_status[npts] receives each point's flag; d_stats (optional)
 * receives [npts][8] counters: nst nfe nje nsetups nni ncfn netf lastflag.
 * d_counter: one zeroed unsigned int of device memory (work queue).
 * precision: PDRCHEM_FP64 or PDRCHEM_MIXED.
 * Returns 0 if the launch was queued. */
int pdrchem_solve_device(int net, double *d_abundance, const double *d_rate, const double *d_density,
                         int npts, double rtol, double atol, double t_start, double t_end, int nelect,
                         int *d_status, int *d_stats, unsigned int *d_counter, int precision, void *stream);

/* As pdrchem_solve_device, allocating nothing -- no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, hosts that forbid hipMalloc).  The
 * caller passes inputs and outputs as for pdrchem_solve_device but no work
 * counter: the solver's scratch memory (work-queue counter and per-block
 * workspace of saved Jacobians and rate constants) is a fixed 48 MiB device
 * array in the library, present on every GPU the library runs on until the
 * process ends, which holds the workspace of a whole-device launch of every
 * network and precision on MI355X.  The point count is not limited by it
 * (the persistent kernel reuses each block's workspace for every point it
 * takes); a launch shape whose single block would not fit is rejected with an
 * error (pdrchem_last_error).  Every call on one GPU shares that GPU's array
 * and overwrites what the last call left there: calls that overlap on one GPU
 * (two streams, two threads) race and must be serialised by the caller.  The
 * counter is cleared by a kernel (no memset node), so the call may be
 * captured into a HIP graph.  Results are those of pdrchem_solve_device.
 * Returns 0 if the launch was queued. */
int pdrchem_solve_device_noalloc(int net, double *d_abundance, const double *d_rate, const double *d_density,
                                 int npts, double rtol, double atol, double t_start, double t_end, int nelect,
                                 int *d_status, int *d_stats, int precision, void *stream);

/* Precision used by calculate_abundances_ (default PDRCHEM_MIXED; the
 * environment variable PDRCHEM_PRECISION=fp64 selects PDRCHEM_FP64). */
void pdrchem_set_precision(int precision);

/* Physical parameters of the rate evaluation (params.dat,
 * input_parameters.F90:19-58): zeta in units of 1.3e-17 s^-1, v_turb in cm/s. */
typedef struct {
  double zeta, v_turb, omega, grain_radius, metallicity, g2d;
} pdrchem_phys;

/* Batched CALCULATE_REACTION_RATES (calc_reac_rates.F90:16-401) on device
 * memory: T, Td [npts]; G, Av [npts*nrays]; cols [npts*nrays*5] column
 * densities of H2, HD, CO, C, S per ray (may be NULL = zero columns);
 * rate [npts*nreac] out; d_neg: one zeroed int that counts points with a
 * negative rate (the reference STOPs on those, calc_reac_rates.F90:387-390).
 * Asynchronous on `stream`. */
int pdrchem_rates_device(int net, const pdrchem_phys *phys, const double *d_T, const double *d_Td,
                         int nrays, const double *d_G, const double *d_Av, const double *d_cols,
                         int npts, double *d_rate, int *d_neg, void *stream);

/* As pdrchem_rates_device, with a cosmic-ray ionisation rate per point:
 * d_zeta[npts] (units of 1.3e-17 s^-1, like phys->zeta; NULL = phys->zeta
 * everywhere) replaces phys->zeta in the cosmic-ray (calc_reac_rates.F90:247),
 * cosmic-ray-photon (:275-276) and cosmic-ray desorption (:334-340) rates.  No
 * counterpart in the reference's rate routine, which takes one global zeta;
 * it serves the per-point zeta of the reference's cosmic-ray fork. */
int pdrchem_rates_device_zeta(int net, const pdrchem_phys *phys, const double *d_T, const double *d_Td,
                              int nrays, const double *d_G, const double *d_Av, const double *d_cols,
                              const double *d_zeta, int npts, double *d_rate, int *d_neg, void *stream);

/* Batched CALC_HEATING (sub_calculate_heating.F90:6-386) on device memory,
 * all three networks (PDRCHEM_NET ids 0-2).  MYNETWORK's chemical heating
 * is the reference's pre-set channel list, which its source guards with a
 * STOP for the user to comment out (sub_calculate_heating.F90:249-260).
 * index[13]: 1-based slots as the Fortran module variables hold them -- NH,
 * NH2, NC, NCO, NHEx, NH3x, NH3Ox, NHCOx, NH2x, NELECT, then reactions NRGR,
 * NRH2, NRCI.  Per point: density, T, Td, uv [npts] (uv in Draine units, as
 * UVfield); abundances [npts*nspec] and rates [npts*nreac] row-major.  Writes
 * heat [npts*12]: photoelectric, PAH photoelectric, Weingartner & Draine, C
 * ionization, H2 formation, H2 photodissociation, FUV pumping, cosmic rays,
 * turbulence, chemical, gas-grain, total -- the order of the reference's
 * .heat.fin columns (3DPDR.F90 output), the total leaving out the
 * photoelectric and Weingartner & Draine terms as the reference's does.
 * Asynchronous on stream; 0 on success, -1 on a bad argument or index. */
int pdrchem_heating_device(int net, const pdrchem_phys *phys, const int *index, int npts,
                           const double *d_density, const double *d_T, const double *d_Td, const double *d_uv,
                           const double *d_abundance, const double *d_rate, double *d_heat, void *stream);

/* As pdrchem_heating_device, with a cosmic-ray ionisation rate per point:
 * d_zeta[npts] (units of 1.3e-17 s^-1, like phys->zeta; NULL = phys->zeta
 * everywhere) in the cosmic-ray heating term (sub_calculate_heating.F90:201),
 * which the reference evaluates with its one global zeta.  For clouds whose
 * rates use a per-point zeta (pdrchem_rates_device_zeta), so that heating and
 * chemistry see the same rate. */
int pdrchem_heating_device_zeta(int net, const pdrchem_phys *phys, const int *index, int npts,
                                const double *d_density, const double *d_T, const double *d_Td, const double *d_uv,
                                const double *d_abundance, const double *d_rate, const double *d_zeta,
                                double *d_heat, void *stream);

/* Solver options (no counterpart in the reference's interface, where they
 * are locals of calculate_abundances.c).  mxstep: internal steps allowed per
 * CVode call before it returns -1 (calculate_abundances.c:111 sets 1e7, the
 * default here); after five such returns a point fails with status -1, and
 * its partial result is written back (:216-305).  Applies to every later
 * solve, request-loop batch and cloud iteration.  Returns 0, -1 if mxstep < 1. */
int pdrchem_set_mxstep(long long mxstep);
long long pdrchem_mxstep(void);

/* What the electron slot holds on return.  PDRCHEM_XE_LAST_RHS (default):
 * x_e of the integrator's last RHS evaluation, as the reference writes it
 * (calculate_abundances.c:317; the charge sum at a Newton iterate of the last
 * step, which may lie past t_end).  PDRCHEM_XE_CHARGE_SUM: the charge sum of
 * the abundances written back, i.e. x_e at t_end.  Returns 0, -1 if unknown. */
#define PDRCHEM_XE_LAST_RHS 0
#define PDRCHEM_XE_CHARGE_SUM 1
int pdrchem_set_electron_policy(int policy);
int pdrchem_electron_policy(void);

/* The solver's Jacobian at given abundances (introspection; no counterpart in
 * the reference, whose analytic jacobian_*.c is not registered,
 * calculate_abundances.c:210-212).  d_abundance [npts*nspec], d_rate
 * [npts*nreac], d_density [npts] as for pdrchem_solve_device; writes d_jac
 * [npts][nspec-1][nspec]: entry (i, j) = d ydot_i / d x_j at fixed x_e for
 * j < nspec-1 and d ydot_i / d x_e for j = nspec-1 (x_e = charge sum of the
 * abundances), evaluated by the device code the solver runs, in the matrix
 * precision of `precision` (PDRCHEM_MIXED rounds each entry to FP32 as the
 * solver stores it).  The Jacobian of CVODE's iteration matrix is
 * J = J0 + g q^T with q the charges.  Asynchronous on stream; 0 = queued. */
int pdrchem_jacobian_device(int net, int precision, const double *d_abundance, const double *d_rate,
                            const double *d_density, int npts, double *d_jac, void *stream);

/* Kernel resource usage of the solver for a network and precision
 * (registers per lane, LDS bytes per block, scratch bytes per lane, blocks
 * per CU).  0 = ok. */
int pdrchem_kernel_info(int net, int precision, int *num_regs, int *lds_bytes, int *scratch_bytes,
                        int *blocks_per_cu);

/* Autotuner of the solver's launch shape (no counterpart in the reference).
 * Runs every candidate shape on the given batch (device memory, as
 * pdrchem_solve_device; d_abundance is restored before each run and on
 * return, so the call leaves it unchanged) and records the fastest for this
 * network, precision and batch-size bucket (point counts in (4^(b-1), 4^b]).
 * Candidates: wavefront-per-system kernel (reduced network): 1, 2 or 4
 * systems per block, and 4 per block with at most 2 blocks per CU;
 * workgroup-per-system kernel: 1 or 2 resident workgroups per CU.  Every
 * later launch of the same network and precision uses the shape recorded for
 * its bucket, or else the one of the largest tuned bucket below it (a device
 * saturated by a few thousand points ranks the shapes the same for any larger
 * batch), or else the default.  Synchronous on `stream`.  Returns 0 and the chosen shape
 * (blocks_per_cu 0 = as many as fit) and its time. */
int pdrchem_autotune(int net, int precision, double *d_abundance, const double *d_rate, const double *d_density,
                     int npts, double t_end_years, int nelect, void *stream, int *systems_per_block,
                     int *blocks_per_cu, double *best_ms);
/* Shape a launch of npts points of this network and precision uses: returns
 * 1 and the shape if a tuned one applies, 0 (and zeros) if the default (4
 * systems per block, full occupancy). */
int pdrchem_tuned_config(int net, int precision, int npts, int *systems_per_block, int *blocks_per_cu);
/* Record a shape for the bucket of npts by hand (a saved tuning restored;
 * systems_per_block 1, 2 or 4, ignored by the workgroup kernel; 0, 0 clears
 * the bucket back to the default). */
int pdrchem_set_tuned_config(int net, int precision, int npts, int systems_per_block, int blocks_per_cu);
/* Forget every tuned shape. */
void pdrchem_clear_tuning(void);
/* Upper bound on resident blocks for the persistent solver (0 = fill the
 * device).  For experiments and tests of the work queue. */
void pdrchem_set_max_blocks(int max_blocks);

/* Text of the last error of this thread (empty if none). */
const char *pdrchem_last_error(void);

/* ---------------------------------------------------------------------
 * Request loop with continuous batching (no counterpart in the reference:
 * for callers that submit many independent solves over time, e.g. an outer
 * iteration that re-solves shrinking, uneven sets of points)
 * ------------------------------------------------------------------- */

typedef struct pdrchem_server pdrchem_server;

/* device_bytes: arena for in-flight batches (0 = 1 GiB; admission is by this
 * footprint); nslots: batches in flight at once, each on its own stream;
 * precision as for pdrchem_solve_device.  Tolerances are read from the
 * globals above at creation.  NULL on failure. */
pdrchem_server *pdrchem_server_create(size_t device_bytes, int nslots, int precision);

/* Queue npts points of network `net` to be integrated from t = 0 to
 * t_end_years with calculate_abundances_ semantics.  abundance[npts*nspec]
 * (host, in/out), rate[npts*nreac] and density[npts] must stay valid until
 * the request is reported complete.  Returns the request's ticket (>= 0) or
 * -1 (bad arguments, request larger than the arena). */
int pdrchem_server_submit(pdrchem_server *server, int net, int npts, double *abundance, const double *rate,
                          const double *density, double t_end_years);

/* As pdrchem_server_submit with device arrays (state resident in HBM): the
 * batch stages them by device-to-device copies and writes the abundances
 * back in place.  d_status[npts] and d_stats[npts][8] (device, optional)
 * receive each point's status and integrator counters as from
 * pdrchem_solve_device.  Device and host requests are never in one batch. */
int pdrchem_server_submit_device(pdrchem_server *server, int net, int npts, double *d_abundance,
                                 const double *d_rate, const double *d_density, double t_end_years, int *d_status,
                                 int *d_stats);
/* Largest number of points of one request (and of one batch); default 65536.
 * Smaller batches trade throughput for latency (a batch's requests complete
 * together). */
int pdrchem_server_max_batch(pdrchem_server *server);
int pdrchem_server_set_max_batch(pdrchem_server *server, int npts);
/* Retire finished batches (results copied into the callers' arrays) and
 * start pending ones; wait != 0 blocks until nothing is pending or in
 * flight.  Returns the number of requests pending or in flight, -1 on error. */
int pdrchem_server_progress(pdrchem_server *server, int wait);

/* Up to max completed requests since the last poll: ticket, status (0, or
 * the CVODE-style flag of its first failed point, or 1 if its batch could
 * not be launched) and submit-to-completion latency in ms.  Returns the count. */
int pdrchem_server_poll(pdrchem_server *server, int max, int *tickets, int *status, double *latency_ms);

/* Standing launches for host requests (pdrchem_server_submit) instead of
 * batches; no counterpart in the reference.  Each network gets a ring of
 * ring_points points (a power of two, 64 .. 2^24; 0 turns standing launches
 * off) in pinned host memory, allocated on the network's first request.  A
 * request is copied into the ring and published; a persistent solver launch
 * keeps taking published points while it runs, so a request starts as soon
 * as a worker is free rather than when a whole batch slot is.  Workers stop
 * after idle_ms without work or life_ms of running; the server keeps a
 * launch running while points wait (a successor is started half way through
 * a launch's life).  Results, statuses and latencies come back through
 * pdrchem_server_progress / pdrchem_server_poll as for batches; a request
 * larger than the ring is refused.  Device requests keep using batches.
 * Refused (-1) while standing launches hold requests; 0 on success. */
int pdrchem_server_set_refill(pdrchem_server *server, int ring_points, double idle_ms, double life_ms);
/* Standing launches started so far and points now held in the rings;
 * returns the ring size (0: standing launches off), -1 on a null server. */
int pdrchem_server_refill_stats(pdrchem_server *server, long long *launches, int *held_points);
/* Batches go to the device as HIP graphs (capture of copy-in, work-queue
 * reset, solver launch and copy-out on the slot's stream; one executable
 * graph per slot, updated in place per batch).  Returns 1 if the server
 * launches graphs (0: plain stream launches, PDRCHEM_SERVER_GRAPHS=0 in the
 * environment at creation) and the graph launches and instantiations so far. */
int pdrchem_server_graph_stats(pdrchem_server *server, long long *launches, long long *instantiations);
const char *pdrchem_server_last_error(pdrchem_server *server);
void pdrchem_server_destroy(pdrchem_server *server);

/* ---------------------------------------------------------------------
 * Grid cloud: ray-traced column densities and the chemistry iteration
 * ------------------------------------------------------------------- */
/* A regular nx*ny*nz grid of points (sides 2..128, spacing in pc; point
 * index ix + nx*(iy + ny*iz)) held on the device with everything the
 * chemistry iteration of 3DPDR.F90:474-520 needs: evaluation points along
 * the 12*4^level HEALPix rays (evaluation_points, eval_points.F90:129-225, in
 * (distance, point index) order, no maxpoints cap), Av per ray and the
 * surface field per ray (CALC_UVFIELD, 3DPDR.F90:1636-1694; fieldchoice 0 =
 * "ISO", 1 = "UNI" with vector gext[3]), the GUESS_TEMP gas temperature
 * (3DPDR.F90:1698-1700, held fixed), dust temperature, abundances, and the
 * columns of H2, HD, CO, C, S per ray (calc_columndens, 3DPDR.F90:1788-1844).
 * shield_slots[5]: 0-based abundance slots of H2, HD, CO, C, S (-1 =
 * absent); nelect 1-based electron slot.  NULL on failure
 * (pdrchem_cloud_last_error(NULL) says why).  No counterpart in the
 * reference as one call; the reference computes the same quantities inside
 * its main program. */
typedef struct pdrchem_cloud pdrchem_cloud;
typedef struct {
  int npts;            /* points in the grid */
  int solved;          /* points solved in this iteration */
  int converged;       /* of those, points whose abundances moved by < conv_tol */
  int failed;          /* of those, points whose solve failed (status != 0) */
  int negative_rates;  /* points with a negative rate coefficient */
  double rates_ms, solve_ms;  /* device time of the rate evaluation and of the solve */
  long long nst, nfe, nje, nsetups, nni;  /* integrator counters summed over the solved points */
} pdrchem_cloud_stats;
pdrchem_cloud *pdrchem_cloud_create(int net, int nx, int ny, int nz, double spacing_pc, const double *density,
                                    int level, double theta_crit, int fieldchoice, const double *gext,
                                    double dust_temperature, const pdrchem_phys *phys, const int *shield_slots,
                                    int nelect);
/* Candidate order of the evaluation-point search among equidistant points */
#define PDRCHEM_ORDER_REFERENCE 0  /* heapsort.F90's order, as the reference runs it */
#define PDRCHEM_ORDER_DISTANCE 1   /* by (distance, point index) */
/* As pdrchem_cloud_create for an arbitrary point cloud (the reference's
 * input: one x, y, z in pc and one n_H per point, read at 3DPDR.F90:166-169):
 * xyz[npts*3]; etype[npts] (NULL = all PDR points; 2 = ionised: a ray stops at
 * it, eval_points.F90:222); order as above.  The evaluation points are found
 * on the host by a threaded restatement of evaluation_points (O(N^2 log N),
 * done once; up to 200,000 points); everything after that runs on the device
 * as for a grid. */
pdrchem_cloud *pdrchem_cloud_create_points(int net, int npts, const double *xyz, const int *etype, int order,
                                           const double *density, int level, double theta_crit, int fieldchoice,
                                           const double *gext, double dust_temperature, const pdrchem_phys *phys,
                                           const int *shield_slots, int nelect);
/* The evaluation points of a point cloud alone (host only, no device use):
 * count[npts*nrays] and offsets[npts*nrays+1] always written, proj and step
 * (projected point ids, step lengths in pc) when capacity >= the total, which
 * is returned (-1 on bad arguments). */
long long pdrchem_trace_points(int npts, const double *xyz, const int *etype, int level, double theta_crit,
                               int order, int *count, long long *offsets, int *proj, double *step,
                               long long capacity);
/* Abundances of every point (host [npts*nspec]); recomputes the columns and
 * marks every point unconverged. */
int pdrchem_cloud_set_abundances(pdrchem_cloud *cloud, const double *abundance);
/* One chemistry iteration (3DPDR.F90:474-520): rates from the current
 * columns, a solve from t = 0 to t_end_years warm-started from the current
 * abundances (calculate_abundances_ semantics, tolerances from the globals
 * above), write-back, new columns.  A point whose solve moved its abundances
 * (those > 1e-20) by less than conv_tol relative is converged and is not
 * solved again: it keeps its abundances, as the reference keeps those of its
 * converged points (3DPDR.F90:628, columns :1822), which it re-solves and discards. */
int pdrchem_cloud_iterate(pdrchem_cloud *cloud, double t_end_years, double conv_tol, int precision,
                          pdrchem_cloud_stats *stats);
/* Cosmic-ray ionisation rate per point (host [npts], units of 1.3e-17 s^-1
 * like phys->zeta; NULL = phys->zeta everywhere) for the rates of later
 * iterations (pdrchem_rates_device_zeta; the per-point zeta of the reference's
 * cosmic-ray fork).  Marks every point unconverged. */
int pdrchem_cloud_set_zeta(pdrchem_cloud *cloud, const double *zeta);
/* Send the iterations' solves through a request loop (device requests of at
 * most pdrchem_server_max_batch points; the loop's precision applies and
 * the precision argument of pdrchem_cloud_iterate is ignored); NULL goes back
 * to direct launches on the cloud's stream.  The server must outlive its use
 * here, and while attached it serves this cloud only: an iteration waits for
 * the loop to drain and clears its completion list (per-point statuses come
 * back through the requests' device arrays). */
int pdrchem_cloud_set_server(pdrchem_cloud *cloud, pdrchem_server *server);
#define PDRCHEM_CLOUD_ABUNDANCE 0    /* [npts*nspec] */
#define PDRCHEM_CLOUD_AV 1           /* [npts*nrays] */
#define PDRCHEM_CLOUD_G 2            /* [npts*nrays] surface field per ray */
#define PDRCHEM_CLOUD_COLUMNS 3      /* [npts*nrays*5] H2, HD, CO, C, S (cm^-2) */
#define PDRCHEM_CLOUD_TEMPERATURE 4  /* [npts] */
#define PDRCHEM_CLOUD_UVFIELD 5      /* [npts] attenuated field */
/* fields of the thermal balance (after pdrchem_cloud_thermal_init) */
#define PDRCHEM_CLOUD_DUST_TEMPERATURE 6    /* [npts] (K) */
#define PDRCHEM_CLOUD_HEATING 7             /* [npts*12] the .heat.fin columns (pdrchem_heating_device) */
#define PDRCHEM_CLOUD_COOLING 8             /* [npts*5] CII, CI, OI, CO line cooling and their sum */
#define PDRCHEM_CLOUD_LINES 9               /* [npts*17] the .line.fin lines (pdrchem_level_populations_device) */
#define PDRCHEM_CLOUD_POPULATIONS 10        /* [npts*nlevtot] CII, CI, OI, CO level populations (cm^-3) */
#define PDRCHEM_CLOUD_PREV_TEMPERATURE 11   /* [npts] previousgastemperature: the reported temperature */
#define PDRCHEM_CLOUD_THERMAL_CONVERGED 12  /* [npts] 1 where the point's thermal balance converged */
#define PDRCHEM_CLOUD_OPTICAL_DEPTHS 13     /* [npts*17] the lines' optical depths along ray 6 (.opdp.fin), or
                                               along the PSEUDO_1D ray */
int pdrchem_cloud_get(pdrchem_cloud *cloud, int field, double *out);
/* Evaluation points: total count, then per (point, ray) the count and CSR
 * offset (offsets[npts*nrays+1]) and per evaluation point the projected
 * point and the step |epoint_i - epoint_{i-1}| in pc (any pointer may be
 * NULL). */
long long pdrchem_cloud_num_evaluation_points(const pdrchem_cloud *cloud);
int pdrchem_cloud_evaluation_points(pdrchem_cloud *cloud, int *count, long long *offsets, int *proj, double *step);

/* ---- thermal balance ---------------------------------------------------- */
/* One coolant's tables as READINPUT leaves them (read_input.F90:6-259), in
 * the reference's column-major layout: X(i, j) of an nlev x nlev matrix at
 * [i + j*nlev]; TEMPERATURES(p, k) at [p + 7k]; col: 7 blocks of
 * nlev*nlev*ntemp, COL(i, j, k) of partner p at [p*nlev*nlev*ntemp + i +
 * j*nlev + k*nlev*nlev].  Partners in the order H2, para-H2, ortho-H2, e-,
 * H, He, H+; zero temperatures mark data a partner lacks.  The coolants
 * are always given as four, CII (5 levels), CI (5), OI (5), CO (up to 41):
 * the reference's order (3DPDR.F90 cooling columns). */
typedef struct {
  int nlev, ntemp;
  const double *energies;      /* [nlev] erg */
  const double *weights;       /* [nlev] statistical weights */
  const double *A, *B, *freq;  /* [nlev*nlev] s^-1, Einstein B, Hz */
  const double *temperatures;  /* [7*ntemp] K */
  const double *col;           /* [7*nlev*nlev*ntemp] cm^3 s^-1 */
} pdrchem_coolant;

/* Device copies of the four coolants' tables and the abundance slots the
 * level populations read: slots[9], 0-based, of C+, C, O, CO (the coolants'
 * species) and of H, H+, e-, He, H2 (the collision partners).  NULL on
 * failure (pdrchem_last_error says why); the level-population calls below
 * report their errors there too. */
typedef struct pdrchem_coolant_tables pdrchem_coolant_tables;
pdrchem_coolant_tables *pdrchem_coolant_tables_create(const pdrchem_coolant *coolants, const int *slots);
void pdrchem_coolant_tables_destroy(pdrchem_coolant_tables *tables);
/* Levels of all four coolants (the row width of population arrays). */
int pdrchem_coolant_tables_levels(const pdrchem_coolant_tables *tables);

/* Batched escape-probability level populations of one coolant (0 CII, 1 CI,
 * 2 OI, 3 CO) on device memory: per point FIND_CCOEFF (find_Ccoeff.F90),
 * escape_probability (escape_probability.F90:6-140) with the optical depth
 * along each of nrays rays summed over the ray's evaluation points, and
 * solvlevpop (solvlevpop.F90:6-56), as the reference's level-population
 * step does (3DPDR.F90:688-850).
 *   d_T, d_Td, d_rho [npts]; d_ab [npts*nspec];
 *   d_pop [nrows*nlevtot] (nrows >= npts): the populations of every level of
 *     every coolant, row p for point p; the evaluation points of ray j of
 *     point p are rows d_proj[d_offs[p*nrays+j] + i], i < d_count[p*nrays+j],
 *     at steps d_step (pc) from the previous one (the first from point p);
 *   pseudo_ray: -1 for 3D (beta the mean over the rays), else the ray of
 *     the PSEUDO_1D build (the other rays optically thick);
 *   phys: v_turb, metallicity, g2d used;
 *   co_fix / co_average_co / levpop_iteration: the CO_FIX averaging with the
 *     old populations (3DPDR.F90:771-846; co_average_co = 1 where the CO
 *     populations are averaged too); co_fix 0 = none;
 *   d_skip [npts] (NULL = none): points left untouched;
 *   d_S: scratch of npts*nrays*nlevtot doubles (the level columns);
 * writes d_solution[npts*nlevtot] (this coolant's levels), d_cool[npts*4]
 * (this coolant's column: line cooling, erg cm^-3 s^-1) and d_lines[npts*17]
 * (this coolant's lines among CII 158um; CI 2-1, 3-1, 3-2; OI 2-1, 3-1, 3-2;
 * CO J=1-0 .. 10-9) and, when d_tau_lines is not NULL, the optical depth
 * of the same lines along ray tau_ray into d_tau_lines[npts*17] (0 for a
 * line without one: empty upper level or |tpop| < 1e-50), which the
 * reference writes for ray 6 to .opdp.fin (3DPDR.F90:1484-1500).  The level
 * columns are formed once per call for all levels; call with coolant = -1 to
 * do all four in turn.  Asynchronous on stream; 0 on success, -1 on a bad
 * argument. */
int pdrchem_level_populations_device(const pdrchem_coolant_tables *tables, int coolant, const pdrchem_phys *phys,
                                     int npts, int nspec, int nrays, int pseudo_ray, const double *d_T,
                                     const double *d_Td, const double *d_rho, const double *d_ab,
                                     const double *d_pop, const int *d_count, const long long *d_offs,
                                     const int *d_proj, const double *d_step, int co_fix, int co_average_co,
                                     int levpop_iteration, const unsigned char *d_skip, double *d_S,
                                     double *d_solution, double *d_cool, double *d_lines, int tau_ray,
                                     double *d_tau_lines, void *stream);
/* LTE populations (3DPDR.F90:520-552) of every coolant at the gas
 * temperature: d_pop[npts*nlevtot]; points with d_skip[p] != 0 untouched. */
int pdrchem_lte_populations_device(const pdrchem_coolant_tables *tables, int npts, int nspec, const double *d_T,
                                   const double *d_ab, const double *d_rho, const unsigned char *d_skip,
                                   double *d_pop, void *stream);
/* CALCULATE_DUST_TEMPERATURES (dust_t.F90:44-97): d_G [npts*nrays] surface
 * field per ray, d_uv [npts] attenuated field -> d_Td [npts]. */
int pdrchem_dust_temperature_device(int npts, int nrays, const double *d_G, const double *d_uv, double *d_Td,
                                    void *stream);

/* Thermal-balance parameters (params.dat, input_parameters.F90:45-51) and
 * the reference's build options. */
typedef struct {
  double Tmin, Tmax;  /* absolute temperature bounds (K) */
  double Fcrit;       /* convergence: 2|H - C|/|H + C| <= Fcrit */
  double Tdiff;       /* bracket width below which the search gives up (K) */
  int pseudo_ray;     /* -1: 3D; else PSEUDO_1D along this ray (6 in the reference) */
  int dust;           /* 1: dust temperatures from the field (DUST) */
  int co_fix;         /* CO_FIX averaging of slow level populations */
  int temp_fix;       /* TEMP_FIX bracket handling */
} pdrchem_thermal_params;
typedef struct {
  int iteration;           /* ITERATION after this call */
  int levpop_iteration;    /* levpop_iteration after this call (0: levels converged) */
  int level_conv;          /* every point's populations converged in this call */
  int first_time;
  int done;                /* every point thermally converged: the loop ends */
  int lvg;                 /* 1: escape-probability populations; 0: chemistry + LTE */
  int thermal_points;      /* points whose heating and temperature were updated */
  int thermally_converged; /* points thermally converged */
  int level_converged;     /* points whose populations are converged */
  int coolant_converged[4];
  int chemistry_solves, failed;
  long long nst, nfe, nje, nsetups, nni;
  double ms;
} pdrchem_thermal_stats;
/* Turn on the thermal balance of a cloud (THERMALBALANCE with GUESS_TEMP):
 * coolant tables and slots as pdrchem_coolant_tables_create; heat_index[13]
 * as pdrchem_heating_device.  Sets the temperature brackets of every point
 * (3DPDR.F90:1702-1716) and, with prm->dust, the dust temperatures from the
 * field.  Works with all three networks. */
int pdrchem_cloud_thermal_init(pdrchem_cloud *cloud, const pdrchem_coolant *coolants, const int *slots,
                               const int *heat_index, const pdrchem_thermal_params *prm);
/* One ITERATION of the reference's outer loop (3DPDR.F90:590-1241): LTE
 * populations before the first; on the first iteration after the level
 * populations converged, three chemistry iterations of the thermally
 * unconverged points (pdrchem_cloud_iterate) and LTE populations; otherwise
 * the escape-probability populations of the points whose populations have
 * not converged; then heating, line cooling and the temperature update of
 * the thermally unconverged points, the convergence counts and the new
 * populations.  stats may be NULL; stats->done = 1 once every point is
 * thermally converged (further calls do nothing). */
int pdrchem_cloud_thermal_iterate(pdrchem_cloud *cloud, double t_end_years, int precision,
                                  pdrchem_thermal_stats *stats);
const char *pdrchem_cloud_last_error(const pdrchem_cloud *cloud);
void pdrchem_cloud_destroy(pdrchem_cloud *cloud);

#ifdef __cplusplus
}
#endif
#endif
