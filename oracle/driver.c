/*
 * calculate_abundances_ on the CPU (calculate_abundances.c:64-354):
 * per-point CVODE-equivalent solve with the reference's output schedule,
 * retry policy and write-back.  OpenMP dynamic schedule over points as the
 * reference (calculate_abundances.c:124-130).  Test infrastructure (see
 * pdr_oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "cvode_bdf.h"
#include "pdr_oracle.h"

#define SECONDS_IN_YEAR 3.1556926e7

/* steps per CVode call (calculate_abundances.c:111) and what the electron
 * slot holds on return: 0 = x_e of the last RHS evaluation (:317), 1 = the
 * charge sum of the written-back abundances.  Set before a call, never
 * during one. */
static long g_mxstep = 10000000L;
static int g_xe_policy = 0;

void pdr_set_mxstep(long mxstep) { g_mxstep = mxstep > 0 ? mxstep : 10000000L; }
void pdr_set_xe_policy(int policy) { g_xe_policy = policy ? 1 : 0; }

typedef struct {
  const pdr_net *net;
  const double *rate;
  double n_H;
  double x_e; /* charge sum of the most recent RHS evaluation */
  int bordered;
} chem_ud;

static int chem_f(double t, const double *y, double *ydot, void *ud) {
  (void)t;
  chem_ud *c = (chem_ud *)ud;
  pdr_rhs(c->net, c->rate, c->n_H, y, ydot, &c->x_e);
  return 0;
}

static int chem_jac(double t, const double *y, double *J, void *ud) {
  (void)t;
  chem_ud *c = (chem_ud *)ud;
  if (c->bordered) pdr_jac_bordered(c->net, c->rate, c->n_H, y, J);
  else pdr_jac(c->net, c->rate, c->n_H, y, J);
  return 0;
}

/* PDR_MODE_HYBRID needs the network's elimination order and the charges */
static bdf_mem *create_for_net(const pdr_net *net, chem_ud *ud, double rtol, double atol,
                               long mxstep, int mode, double *qbuf) {
  if ((mode & PDR_MODE_HYBRID) && (!net->border_order || net->nb != net->neq + 1)) return NULL;
  bdf_mem *m = bdf_create(net->neq, chem_f, chem_jac, ud, rtol, atol, mxstep, mode);
  if (!m) return NULL;
  if (mode & PDR_MODE_HYBRID) {
    ud->bordered = 1;
    for (int j = 0; j < net->neq; j++) qbuf[j] = net->charge[j];
    bdf_set_border(m, qbuf, net->nhead, net->border_order);
  }
  return m;
}

static void fill_stats(pdr_cvode_stats *st, const bdf_mem *m, int flag) {
  if (!st) return;
  st->nst = m->nst; st->nfe = m->nfe; st->nfeDQ = m->nfeDQ; st->nsetups = m->nsetups;
  st->nje = m->nje; st->nni = m->nni; st->ncfn = m->ncfn; st->netf = m->netf;
  st->qlast = m->qu; st->hlast = m->hu; st->flag = flag; st->nfails = 0;
}

/* one grid point: returns the status contribution of calculate_abundances.c */
static int solve_point(const pdr_net *net, double *ab, const double *rate, double n_H,
                       double reltol, double abstol, double start_time, double end_time,
                       int nelect0, int mode, pdr_cvode_stats *st) {
  const int neq = net->neq;
  const long mxstep = g_mxstep;
  chem_ud ud = {net, rate, n_H, ab[nelect0], 0};
  double *qbuf = malloc(sizeof(double) * (neq + 1));
  bdf_mem *m = qbuf ? create_for_net(net, &ud, reltol, abstol, mxstep, mode, qbuf) : NULL;
  if (!m) { free(qbuf); return 1; }
  double t0 = start_time * SECONDS_IN_YEAR;
  double tend = end_time * SECONDS_IN_YEAR;
  bdf_init(m, t0, ab);
  double *y = malloc(sizeof(double) * neq);
  memcpy(y, ab, sizeof(double) * neq);
  double tout = 1.0e-4 * SECONDS_IN_YEAR;
  double t = t0;
  int nfails = 0, status = 0, flag = 0;
  do {
    flag = bdf_solve(m, tout, y, &t);
    if (flag != 0) {
      nfails++;
      if (flag == PDR_CV_TOO_MUCH_WORK) {
        if (nfails >= 5) status = flag;  /* mxstep re-set to the same value */
      } else if (flag == PDR_CV_CONV_FAILURE) {
        if (nfails >= 5) status = flag;
      } else {
        status = flag;
        nfails = 5;
      }
    }
    tout = tout * 10;
    if (tout > tend) tout = tend;
  } while (t < tout && nfails < 5);
  double xe_sum = 0.0;
  for (int i = 0; i < neq; i++) {
    ab[i] = y[i];
    if (ab[i] < abstol) ab[i] = 0;
    xe_sum += net->charge[i] * ab[i];
  }
  ab[nelect0] = g_xe_policy ? xe_sum : ud.x_e;
  fill_stats(st, m, flag);
  if (st) st->nfails = nfails;
  free(y);
  bdf_free(m);
  free(qbuf);
  return status;
}

int pdr_calculate_abundances(double *abundance, const double *rate, const double *density,
                             const double *temperature, int npart, int nspec, int nreac,
                             const pdr_net *net, double reltol, double abstol,
                             double start_time, double end_time, int nelect, int mode,
                             int nthreads, pdr_cvode_stats *stats) {
  (void)temperature;
  if (!net || nspec != net->nspec || nreac != net->nreac) return 1;
  int status = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
#endif
  for (int n = 0; n < npart; n++) {
    int s = solve_point(net, abundance + (size_t)n * nspec, rate + (size_t)n * nreac, density[n],
                        reltol, abstol, start_time, end_time, nelect - 1, mode,
                        stats ? stats + n : NULL);
    if (s != 0) {
#pragma omp critical(pdr_status)
      status = s;
    }
  }
  return status;
}

int pdr_integrate_point(const pdr_net *net, const double *rate, double n_H, const double *y0,
                        double t0, int ntout, const double *tout, double *yout,
                        const pdr_cvode_opts *o, pdr_cvode_stats *st) {
  chem_ud ud = {net, rate, n_H, 0.0, 0};
  double *qbuf = malloc(sizeof(double) * (net->neq + 1));
  bdf_mem *m = qbuf ? create_for_net(net, &ud, o->rtol, o->atol, o->mxstep, o->mode, qbuf) : NULL;
  if (!m) { free(qbuf); return 1; }
  bdf_init(m, t0, y0);
  int flag = 0;
  for (int k = 0; k < ntout; k++) {
    double t;
    flag = bdf_solve(m, tout[k], yout + (size_t)k * net->neq, &t);
    if (flag != 0) break;
  }
  fill_stats(st, m, flag);
  bdf_free(m);
  free(qbuf);
  return flag;
}

int pdr_integrate_generic(int n, pdr_rhs_fn f, pdr_jac_fn jac, void *ud, const double *y0,
                          double t0, int ntout, const double *tout, double *yout,
                          const pdr_cvode_opts *o, pdr_cvode_stats *st) {
  bdf_mem *m = bdf_create(n, f, jac, ud, o->rtol, o->atol, o->mxstep, o->mode);
  if (!m) return 1;
  bdf_init(m, t0, y0);
  int flag = 0;
  for (int k = 0; k < ntout; k++) {
    double t;
    flag = bdf_solve(m, tout[k], yout + (size_t)k * n, &t);
    if (flag != 0) break;
  }
  fill_stats(st, m, flag);
  bdf_free(m);
  return flag;
}
