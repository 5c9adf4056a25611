/*
 * pdr_oracle.h -- CPU restatement of 3D-PDR's per-grid-point chemistry path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product path: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library never links it.
 *
 * Contents and the reference code each part restates:
 *   pdr_net_*        network tables written by the network compiler
 *                    (read_species.F90:91-190, read_rates.F90:53-91)
 *   pdr_rates        CALCULATE_REACTION_RATES  calc_reac_rates.F90:16-401
 *                    with H2PDRATE/COPDRATE/CIPDRATE/SIPDRATE, H2SHIELD1,
 *                    COSHIELD, SCATTER, XLAMBDA, LBAR (shield.F90),
 *                    H2_FORMATION_RATE (h2_form.F90:8-113) and natural cubic
 *                    splines (spline.F90, restated from the mathematics)
 *   pdr_rhs          ydot = f(x), the MakeRates RHS of odes_*.c:26-148
 *   pdr_jac          exact dense Jacobian (jacobian_*.c plus the x_e chain
 *                    terms the shipped file omits)
 *   pdr_cvode_*      CVODE 2.7.0 (SUNDIALS 2.5.0) BDF + Newton + dense LU,
 *                    restated from its published algorithm (Hindmarsh et al.
 *                    2005, ACM TOMS 31:363; CVODE user guide, "Mathematical
 *                    considerations"); the library itself is absent from the
 *                    reference tree (.MISSING_LARGE_BLOBS:1), so its parity is
 *                    unpinned by the reference: no reference output or test
 *                    exercises the integrator.  It is pinned instead by closed
 *                    forms (linear decay), Robertson's problem, Radau (scipy,
 *                    rtol 1e-12) solutions of slab points and element/charge
 *                    conservation (tests/test_oracle.py)
 *   pdr_calculate_abundances   calculate_abundances_ calculate_abundances.c:64-354
 *   pdr_eval_points, pdr_ray_columns, pdr_uvfield, pdr_hp_*
 *                    evaluation points, Av and column densities along the
 *                    HEALPix rays (eval_points.F90:129-225, healpix.F90,
 *                    3DPDR.F90:1636-1708,1818-1838), in raytrace.c
 */
#ifndef PDR_ORACLE_H
#define PDR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PT_THERMAL = 0, PT_H2FORM, PT_PAH, PT_PHOTO, PT_PHOTO_H2, PT_PHOTO_HD,
  PT_PHOTO_CO, PT_PHOTO_CI, PT_PHOTO_SI, PT_CRP, PT_CRPHOT, PT_FREEZE,
  PT_ELFRZE, PT_CRH, PT_PHOTD, PT_THERM, PT_GRAIN, PT_XR
};

typedef struct {
  int nspec, nreac, neq;
  int xlambda_spline;   /* 1: XLAMBDA uses its cubic spline; 0: linear (see netgen) */
  int *charge;          /* [nspec] */
  double *abund0;       /* [nspec] species-file initial abundances */
  double *mass;         /* [nspec] */
  int *rtype, *dup, *gflag, *fa, *fb, *nh; /* [nreac] */
  double *alpha, *beta, *gamma, *rtmin, *rtmax; /* [nreac] */
  int *st_ptr;          /* [nreac+1] CSR of net stoichiometry */
  int *st_idx, *st_coef;
  /* bordered iteration matrix (unknowns x_0..x_{neq-1}, x_e): size nb =
   * neq+1, static elimination order, first nhead pivots eliminated without
   * pivoting, the rest inverted by Gauss-Jordan (pdrchem/sparse.py) */
  int nb, nhead;
  int *border_order;    /* [nb] */
} pdr_net;

int pdr_net_load(const char *path, pdr_net *net);
void pdr_net_free(pdr_net *net);
pdr_net *pdr_net_open(const char *path); /* heap handle for ctypes */
void pdr_net_close(pdr_net *net);
int pdr_net_nspec(const pdr_net *net);
int pdr_net_nreac(const pdr_net *net);

/* Global physical parameters (params.dat, input_parameters.F90:19-58) */
typedef struct {
  double zeta;          /* cosmic-ray rate in units of 1.3e-17 s^-1 (input_parameters.F90:31) */
  double v_turb;        /* turbulent velocity, cm/s (3DPDR.F90:47) */
  double omega;         /* dust albedo */
  double grain_radius;  /* as read from params.dat */
  double metallicity;
  double g2d;           /* gas-to-dust ratio */
} pdr_phys;

/* Per-point inputs of CALCULATE_REACTION_RATES.  cols[k*5 + c] holds the
 * column density along ray k of H2, HD, CO, C, S (c = 0..4). */
int pdr_rates(const pdr_net *net, const pdr_phys *ph, double T, double Td,
              int nrays, const double *G, const double *Av, const double *cols,
              double *rate);

/* Batched: arrays are point-major; returns number of points with a negative
 * rate (the reference STOPs on those, calc_reac_rates.F90:387-390). */
int pdr_rates_batch(const pdr_net *net, const pdr_phys *ph, int npts,
                    const double *T, const double *Td, int nrays,
                    const double *G, const double *Av, const double *cols,
                    double *rate);

void pdr_rhs(const pdr_net *net, const double *rate, double n_H,
             const double *x, double *ydot, double *x_e);
void pdr_jac(const pdr_net *net, const double *rate, double n_H,
             const double *x, double *J /* row-major neq*neq */);
/* Jacobian of the bordered system: Jb row-major (neq+1)^2, rows i < neq hold
 * d ydot_i / d x_j at fixed x_e (columns j < neq) and d ydot_i / d x_e
 * (column neq); row neq is zero */
void pdr_jac_bordered(const pdr_net *net, const double *rate, double n_H,
                      const double *x, double *Jb);

/* integrator method bits.  0 is the reference configuration: CVODE's
 * difference-quotient Jacobian and a plain dense LU.  The HIP solver runs
 * PDR_MODE_GPU: exact Jacobian, iteration matrix scaled by the error weights
 * (row i by w_i, column j by 1/w_j) and inverted explicitly by Gauss-Jordan
 * with partial pivoting; PDR_MODE_GPU_MIXED also rounds the saved J and the
 * inverse to FP32.  REFINE(k) adds k steps of iterative refinement with an
 * FP64 residual (used to compare linear-solve variants). */
#define PDR_MODE_EXACT_JAC 1
#define PDR_MODE_SCALED 2
#define PDR_MODE_REFINE(k) ((k) << 2) /* k = 0..3 refinement steps */
#define PDR_MODE_F32 16               /* saved J and factorisation rounded to FP32 */
#define PDR_MODE_GJ 64                /* explicit Gauss-Jordan inverse instead of LU */
#define PDR_MODE_STATIC 128           /* LU without pivoting in a fixed order */
#define PDR_MODE_HYBRID 256           /* bordered matrix: static head + Gauss-Jordan tail */
#define PDR_MODE_GPU_WG (PDR_MODE_EXACT_JAC | PDR_MODE_SCALED | PDR_MODE_HYBRID)
#define PDR_MODE_GPU_WG_MIXED (PDR_MODE_GPU_WG | PDR_MODE_F32)
#define PDR_MODE_GPU (PDR_MODE_EXACT_JAC | PDR_MODE_SCALED | PDR_MODE_GJ)
#define PDR_MODE_GPU_MIXED (PDR_MODE_GPU | PDR_MODE_F32)

/* elimination order used by PDR_MODE_STATIC (n <= 512) */
void pdr_set_static_order(int n, const int *order);

/* integrator options / statistics */
typedef struct {
  double rtol, atol;
  long mxstep;
  int mode;             /* PDR_MODE_* bits; 0 = the reference configuration */
  int maxord;           /* 5 */
} pdr_cvode_opts;

typedef struct {
  long nst, nfe, nfeDQ, nsetups, nje, nni, ncfn, netf;
  int qlast;
  double hlast;
  int flag;             /* last CVode return flag */
  int nfails;           /* CVode calls of pdr_calculate_abundances that failed (retried or final) */
} pdr_cvode_stats;

/* calculate_abundances_ (calculate_abundances.c:64-354) on the CPU:
 * abundance[npart*nspec] in/out, rate[npart*nreac], density[npart],
 * temperature[npart]; start/end time in years; nelect 1-based.
 * stats (optional) receives per-point counters.  nthreads<=0: OpenMP default. */
int pdr_calculate_abundances(double *abundance, const double *rate,
                             const double *density, const double *temperature,
                             int npart, int nspec, int nreac,
                             const pdr_net *net, double reltol, double abstol,
                             double start_time, double end_time, int nelect,
                             int mode, int nthreads, pdr_cvode_stats *stats);

/* Options of pdr_calculate_abundances (set before a call, not during one):
 * steps allowed per CVode call (calculate_abundances.c:111, default 1e7),
 * and the electron slot on return: 0 = x_e of the last RHS evaluation
 * (:317, default), 1 = the charge sum of the written-back abundances. */
void pdr_set_mxstep(long mxstep);
void pdr_set_xe_policy(int policy);

/* Status of the bordered head/tail factorisation of one nb x nb matrix
 * (PDR_MODE_HYBRID's setup): 0, or 1 if a head pivot is zero or not finite
 * (in FP32, f32 != 0, also if its reciprocal is not a finite normal float)
 * or the tail is singular.  a is overwritten.  For tests of the rule. */
int pdr_test_hybrid_factor(int nb, double *a, const int *order, int nhead, int f32);

/* Single integration of one point to a list of output times (seconds);
 * writes y at each time into yout[ntout*neq].  For integrator unit tests. */
int pdr_integrate_point(const pdr_net *net, const double *rate, double n_H,
                        const double *y0, double t0, int ntout,
                        const double *tout, double *yout,
                        const pdr_cvode_opts *opts, pdr_cvode_stats *st);

/* Generic ODE integration with the same integrator (for closed-form tests) */
typedef int (*pdr_rhs_fn)(double t, const double *y, double *ydot, void *ud);
typedef int (*pdr_jac_fn)(double t, const double *y, double *J, void *ud);
int pdr_integrate_generic(int n, pdr_rhs_fn f, pdr_jac_fn jac, void *ud,
                          const double *y0, double t0, int ntout,
                          const double *tout, double *yout,
                          const pdr_cvode_opts *opts, pdr_cvode_stats *st);

/* heating terms of a point (heating.c): CALC_HEATING, sub_calculate_heating.F90 */
void pdr_heating(const pdr_phys *ph, int net, const int *idx, double n, double T, double Td, double uv,
                 const double *ab, const double *rate, double *heat);

/* level populations and line cooling (thermal.c).  A coolant's tables in
 * the reference's column-major layout (read_input.F90): X(i, j) of an
 * nlev x nlev matrix at [i + j*nlev] (0-based), TEMPERATURES(p, k) at
 * [p + 7k], partner p's rate COL(i, j, k) at col[p*nlev*nlev*ntemp + i +
 * j*nlev + k*nlev*nlev]; partners H2, pH2, oH2, e, H, He, H+. */
typedef struct {
  int nlev, ntemp;
  const double *E, *g, *A, *B, *F;
  const double *temps, *col;
} pdr_coolant;
/* FIND_CCOEFF (find_Ccoeff.F90): dens = n(H), n(H+), n(e), n(He), n(H2) */
void pdr_find_ccoeff(const pdr_coolant *c, double T, const double *dens, double *C);
/* escape_probability (escape_probability.F90); see thermal.c */
double pdr_escape_probability(const pdr_coolant *c, double Td, double Tg, double v_turb, double density,
                              double metallicity, double g2d, int nrays, int pseudo_ray, const int *epray,
                              const long long *pop_off, const double *chain_pop, const long long *step_off,
                              const double *step, const double *pop, const double *C, double *transition,
                              double *line, double *tau, double *beta);
/* solvlevpop (solvlevpop.F90); returns 1 on a singular matrix */
int pdr_solvlevpop(int n, const double *transition, double density, double *solution);
/* LTE populations (3DPDR.F90:1848-1898) */
void pdr_lte_populations(const pdr_coolant *c, double T, double density, double *pop);
/* CALCULATE_DUST_TEMPERATURES for one point (dust_t.F90:44-97) */
double pdr_dust_temperature(double uv, int nrays, const double *rad_surface);

/* ray tracer (raytrace.c) */
#define PDR_ORDER_HEAPSORT 0  /* the reference's heapsort order (ties as it leaves them) */
#define PDR_ORDER_STABLE 1    /* strict (distance, point index) order */
void pdr_hp_pix2vec_nest(int nside, int ipix, double v[3]);
int pdr_hp_vec2pix_nest(int nside, const double r[3]);
int pdr_eval_points(int npts, const double *xyz, const int *etype, int level, double theta_crit, int order,
                    int maxpoints, int *epray, int *proj, double *epoint);
void pdr_ray_columns(int npts, int nrays, int maxpoints, const int *epray, const int *proj,
                     const double *epoint, const double *rho, int nsel, const double *xsel, double av_fac,
                     double *Av, double *cols);
void pdr_uvfield(int npts, int level, int fieldchoice, const double gext[3], const double *Av, double uv_fac,
                 double *G, double *uv, double *tguess);

#ifdef __cplusplus
}
#endif
#endif
