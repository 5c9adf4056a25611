/*
 * Internal state of the CVODE-equivalent BDF integrator (cvode_bdf.c).
 * Test infrastructure (see pdr_oracle.h).
 */
#ifndef PDR_CVODE_BDF_H
#define PDR_CVODE_BDF_H

#include "pdr_oracle.h"

/* return flags, numerically equal to CVODE's (cvode.h) */
#define PDR_CV_SUCCESS 0
#define PDR_CV_TOO_MUCH_WORK -1
#define PDR_CV_TOO_MUCH_ACC -2
#define PDR_CV_ERR_FAILURE -3
#define PDR_CV_CONV_FAILURE -4
#define PDR_CV_LSETUP_FAIL -6
#define PDR_CV_RHSFUNC_FAIL -8
#define PDR_CV_FIRST_RHSFUNC_ERR -9
#define PDR_CV_REPTD_RHSFUNC_ERR -10
#define PDR_CV_UNREC_RHSFUNC_ERR -11
#define PDR_CV_BAD_T -25
#define PDR_CV_TOO_CLOSE -27

typedef struct {
  int n;
  pdr_rhs_fn f;
  pdr_jac_fn jac;
  void *ud;
  double rtol, atol;
  long mxstep;
  int mode;
  int qmax;
  /* Nordsieck array and work vectors */
  double *zn[6];
  double *ewt, *acor, *y, *tempv, *ftemp, *vtmp;
  double *M, *savedJ;
  double *wsc;          /* error weights frozen at the last factorisation */
  double *rsd;          /* refinement work vector */
  double *gjw;          /* 2n work vector of the Gauss-Jordan solve */
  int *piv;
  double *buf;
  /* PDR_MODE_HYBRID: bordered matrix of size nb = n + 1 (last unknown x_e) */
  int nb, nhead;
  const double *bq;         /* [n] charges: x_e = sum_j bq_j x_j */
  const int *border_order;  /* [nb] elimination order */
  double *tail;             /* inverse of the Schur complement, (nb-nhead)^2 */
  int *tpiv;
  double *bwork;            /* 4*nb */
  /* step state */
  double tn, h, hscale, hprime, hu, eta, etamax, hmin, hmax_inv;
  double tau[7], tq[6], l[6];
  double rl1, gamma, gammap, gamrat, crate, acnrm, saved_tq5;
  int q, qprime, qu, L, qwait, indx_acor, nscon, jcur, first;
  long nst, nfe, nfeDQ, nsetups, nje, nni, ncfn, netf, nstlp, nstlj;
} bdf_mem;

bdf_mem *bdf_create(int n, pdr_rhs_fn f, pdr_jac_fn jac, void *ud, double rtol, double atol,
                    long mxstep, int mode);
void bdf_free(bdf_mem *m);
/* PDR_MODE_HYBRID only: border charges and the static elimination order */
void bdf_set_border(bdf_mem *m, const double *q, int nhead, const int *order);
int bdf_init(bdf_mem *m, double t0, const double *y0);
int bdf_solve(bdf_mem *m, double tout, double *yout, double *tret);
int bdf_get_dky(bdf_mem *m, double t, double *dky);

#endif
