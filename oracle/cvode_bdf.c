/*
 * CVODE-equivalent variable-order BDF integrator (FP64).
 * Test infrastructure (see pdr_oracle.h).
 *
 * The reference links SUNDIALS CVODE 2.7.0 (SUNDIALS 2.5.0; makefile:8-10,
 * call sites calculate_abundances.c:148-218) with CV_BDF + CV_NEWTON,
 * CVDense and the library's difference-quotient Jacobian
 * (calculate_abundances.c:181,207,210-212).  The library is not in the
 * reference tree (.MISSING_LARGE_BLOBS:1).  This file restates its published
 * algorithm: Nordsieck history array, fixed-leading-coefficient BDF of
 * orders 1-5 (Jackson & Sacks-Davis 1980), local error test on the WRMS norm
 * of the corrector increment, step/order selection with the bias factors
 * 6/6/10 and the 1.5 threshold, Newton iteration (at most 3 corrections,
 * convergence constant 0.1, rate estimate with 0.3 decay, divergence at
 * ratio 2), iteration-matrix refresh every 20 steps or when gamma moves by
 * more than 30 %, Jacobian refresh every 50 steps or on failures, at most 7
 * error-test and 10 convergence failures per step, the h0 estimate from a
 * trial forward-Euler step, and CV_NORMAL output by Nordsieck interpolation
 * (Hindmarsh et al. 2005, ACM TOMS 31:363; CVODE user guide v2.7.0).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pdr_oracle.h"
#include "cvode_bdf.h"

#define QMAX 5
#define UROUND DBL_EPSILON

/* algorithm constants (CVODE user guide, optional-input defaults and the
 * step-control description) */
#define ETAMX1 10000.0
#define ETAMX2 10.0
#define ETAMX3 10.0
#define ETAMXF 0.2
#define ETAMIN 0.1
#define ETACF 0.25
#define ADDON 0.000001
#define BIAS1 6.0
#define BIAS2 6.0
#define BIAS3 10.0
#define ONEPSM 1.000001
#define THRESH 1.5
#define SMALL_NST 10
#define MXNCF 10
#define MXNEF 7
#define MXNEF1 3
#define SMALL_NEF 2
#define LONG_WAIT 10
#define NLS_MAXCOR 3
#define CRDOWN 0.3
#define DGMAX 0.3
#define RDIV 2.0
#define MSBP 20
#define MSBJ 50
#define DGMAX_J 0.2
#define NLSCOEF 0.1
#define HLB_FACTOR 100.0
#define HUB_FACTOR 0.1
#define H_BIAS 0.5
#define MAX_ITERS 4
#define FUZZ_FACTOR 100.0
#define MIN_INC_MULT 1000.0

/* internal flags */
enum { DO_ERROR_TEST = 2, PREDICT_AGAIN = 3, CONV_FAIL = 4, TRY_AGAIN = 5,
       FIRST_CALL = 6, PREV_CONV_FAIL = 7, PREV_ERR_FAIL = 8, RHSFUNC_RECVR = 9 };
enum { CV_NO_FAILURES = 0, CV_FAIL_BAD_J = 1, CV_FAIL_OTHER = 2 };

static double wrms(int n, const double *v, const double *w) {
  double s = 0.0;
  for (int i = 0; i < n; i++) { double p = v[i] * w[i]; s += p * p; }
  return sqrt(s / n);
}

static void set_ewt(bdf_mem *m, const double *y) {
  for (int i = 0; i < m->n; i++) m->ewt[i] = 1.0 / (m->rtol * fabs(y[i]) + m->atol);
}

static int call_f(bdf_mem *m, double t, const double *y, double *yd) {
  return m->f(t, y, yd, m->ud);
}

bdf_mem *bdf_create(int n, pdr_rhs_fn f, pdr_jac_fn jac, void *ud, double rtol, double atol,
                    long mxstep, int mode) {
  bdf_mem *m = calloc(1, sizeof(bdf_mem));
  if (!m) return NULL;
  m->n = n; m->f = f; m->jac = jac; m->ud = ud;
  m->rtol = rtol; m->atol = atol; m->mxstep = mxstep; m->mode = mode;
  m->qmax = QMAX;
  m->nb = (mode & PDR_MODE_HYBRID) ? n + 1 : n;
  size_t v = (size_t)m->nb;
  m->buf = calloc(v * (QMAX + 1 + 10) + 3 * v * v + 4 * v, sizeof(double));
  m->piv = calloc(2 * v, sizeof(int));
  if (!m->buf || !m->piv) { bdf_free(m); return NULL; }
  m->tpiv = m->piv + v;
  for (int j = 0; j <= QMAX; j++) m->zn[j] = m->buf + v * j;
  double *p = m->buf + v * (QMAX + 1);
  m->ewt = p; p += v;
  m->acor = p; p += v;
  m->y = p; p += v;
  m->tempv = p; p += v;
  m->ftemp = p; p += v;
  m->vtmp = p; p += v;
  m->wsc = p; p += v;
  m->rsd = p; p += v;
  m->gjw = p; p += 2 * v;
  m->M = p; p += v * v;
  m->savedJ = p; p += v * v;
  m->tail = p; p += v * v;
  m->bwork = p;
  return m;
}

void bdf_free(bdf_mem *m) {
  if (!m) return;
  free(m->buf);
  free(m->piv);
  free(m);
}

void bdf_set_border(bdf_mem *m, const double *q, int nhead, const int *order) {
  m->bq = q;
  m->nhead = nhead;
  m->border_order = order;
}

int bdf_init(bdf_mem *m, double t0, const double *y0) {
  memcpy(m->zn[0], y0, sizeof(double) * m->n);
  m->tn = t0;
  m->q = 1; m->L = 2; m->qwait = m->L;
  m->etamax = ETAMX1;
  m->qu = 0; m->hu = 0.0; m->h = 0.0;
  m->nst = m->nfe = m->nfeDQ = m->nsetups = m->nje = m->nni = m->ncfn = m->netf = 0;
  m->nstlp = m->nstlj = 0;
  m->saved_tq5 = 0.0;
  m->indx_acor = QMAX;
  for (int i = 0; i <= QMAX + 1; i++) m->tau[i] = 0.0;
  for (int i = 0; i <= 5; i++) m->tq[i] = 0.0;
  m->gammap = m->gamma = m->gamrat = 0.0;
  m->crate = 1.0;
  m->first = 1;
  return 0;
}

/* ---------------- dense LU with partial pivoting (row-major) ---------- */
/* f32 != 0: every factor entry is rounded to FP32 as it is formed, which
 * emulates a factorisation carried out in single precision */
static int lu_factor(int n, double *a, int *piv, int f32) {
  if (f32)
    for (size_t i = 0; i < (size_t)n * n; i++) a[i] = (float)a[i];
  for (int k = 0; k < n; k++) {
    int p = k;
    double mx = fabs(a[(size_t)k * n + k]);
    for (int i = k + 1; i < n; i++) {
      double v = fabs(a[(size_t)i * n + k]);
      if (v > mx) { mx = v; p = i; }
    }
    piv[k] = p;
    if (a[(size_t)p * n + k] == 0.0) return k + 1;
    if (p != k)
      for (int j = 0; j < n; j++) {
        double t = a[(size_t)k * n + j];
        a[(size_t)k * n + j] = a[(size_t)p * n + j];
        a[(size_t)p * n + j] = t;
      }
    double inv = 1.0 / a[(size_t)k * n + k];
    for (int i = k + 1; i < n; i++) {
      double l = a[(size_t)i * n + k] * inv;
      if (f32) l = (float)l;
      a[(size_t)i * n + k] = l;
      if (l != 0.0)
        for (int j = k + 1; j < n; j++) {
          a[(size_t)i * n + j] -= l * a[(size_t)k * n + j];
          if (f32) a[(size_t)i * n + j] = (float)a[(size_t)i * n + j];
        }
    }
  }
  return 0;
}

/* In-place Gauss-Jordan inverse with partial pivoting, in the order the HIP
 * kernel uses: at step k the pivot is the row, among those not yet used,
 * with the largest |a[i][k]| (first such row on ties); rows are not swapped,
 * piv[k] records the pivot row.  On return a holds the inverse in pivot
 * order, applied by gj_apply.  f32 rounds every entry to FP32 as formed.
 * Returns 1 if the matrix is singular. */
static int gj_invert(int n, double *a, int *piv, int f32) {
  int used[512];
  if (n > 512) return 1;
  for (int i = 0; i < n; i++) {
    used[i] = 0;
    if (f32)
      for (int j = 0; j < n; j++) a[(size_t)i * n + j] = (float)a[(size_t)i * n + j];
  }
  for (int k = 0; k < n; k++) {
    int p = -1;
    double vm = -1.0;
    for (int i = 0; i < n; i++)
      if (!used[i]) {
        double v = fabs(a[(size_t)i * n + k]);
        if (v > vm) { vm = v; p = i; }
      }
    if (!(vm > 0.0)) return 1;
    used[p] = 1;
    piv[k] = p;
    double d = 1.0 / a[(size_t)p * n + k];
    if (f32) d = (float)d;
    double *rp = a + (size_t)p * n;
    for (int j = 0; j < n; j++) {
      double v = (j == k) ? d : rp[j] * d;
      rp[j] = f32 ? (float)v : v;
    }
    for (int i = 0; i < n; i++) {
      if (i == p) continue;
      double *ri = a + (size_t)i * n;
      double fi = ri[k];
      for (int j = 0; j < n; j++) {
        double v = (j == k) ? -fi * d : ri[j] - fi * rp[j];
        ri[j] = f32 ? (float)v : v;
      }
    }
  }
  return 0;
}

/* LU without pivoting in a fixed elimination order (the order a sparse
 * factorisation computes once per network): step k eliminates with the
 * diagonal entry of row/column order[k].  piv[] receives order[].  Returns
 * k+1 if the k-th pivot is zero. */
static int g_static_n = 0;
static int g_static_order[512];

void pdr_set_static_order(int n, const int *order) {
  g_static_n = (n > 0 && n <= 512) ? n : 0;
  for (int i = 0; i < g_static_n; i++) g_static_order[i] = order[i];
}

static int lu_factor_static(int n, double *a, int *piv, int f32) {
  if (g_static_n != n) return 1;
  for (int k = 0; k < n; k++) piv[k] = g_static_order[k];
  if (f32)
    for (size_t i = 0; i < (size_t)n * n; i++) a[i] = (float)a[i];
  for (int k = 0; k < n; k++) {
    const int pk = piv[k];
    const double d = a[(size_t)pk * n + pk];
    if (d == 0.0) return k + 1;
    const double inv = 1.0 / d;
    for (int kk = k + 1; kk < n; kk++) {
      const int i = piv[kk];
      double l = a[(size_t)i * n + pk] * inv;
      if (f32) l = (float)l;
      a[(size_t)i * n + pk] = l;
      if (l == 0.0) continue;
      for (int jj = k + 1; jj < n; jj++) {
        const int j = piv[jj];
        double v = a[(size_t)i * n + j] - l * a[(size_t)pk * n + j];
        a[(size_t)i * n + j] = f32 ? (float)v : v;
      }
    }
  }
  return 0;
}

static void lu_solve_static(int n, const double *a, const int *piv, double *b) {
  for (int k = 0; k < n; k++) {
    const int pk = piv[k];
    for (int kk = k + 1; kk < n; kk++) {
      const int i = piv[kk];
      b[i] -= a[(size_t)i * n + pk] * b[pk];
    }
  }
  for (int k = n - 1; k >= 0; k--) {
    const int pk = piv[k];
    b[pk] /= a[(size_t)pk * n + pk];
    for (int kk = 0; kk < k; kk++) {
      const int i = piv[kk];
      b[i] -= a[(size_t)i * n + pk] * b[pk];
    }
  }
}

/* x = M^-1 b with the pivot-ordered inverse: x_k = sum_j A[p_k][j] b[p_j] */
static void gj_apply(int n, const double *a, const int *piv, double *b, double *work) {
  double *bp = work, *y = work + n;
  for (int j = 0; j < n; j++) bp[j] = b[piv[j]];
  for (int i = 0; i < n; i++) {
    double acc = 0.0;
    for (int j = 0; j < n; j++) acc += a[(size_t)i * n + j] * bp[j];
    y[i] = acc;
  }
  for (int k = 0; k < n; k++) b[k] = y[piv[k]];
}

static void lu_solve(int n, const double *a, const int *piv, double *b) {
  for (int k = 0; k < n; k++) {
    int p = piv[k];
    if (p != k) { double t = b[k]; b[k] = b[p]; b[p] = t; }
    for (int i = k + 1; i < n; i++) b[i] -= a[(size_t)i * n + k] * b[k];
  }
  for (int k = n - 1; k >= 0; k--) {
    b[k] /= a[(size_t)k * n + k];
    for (int i = 0; i < k; i++) b[i] -= a[(size_t)i * n + k] * b[k];
  }
}

/* ---------------- bordered hybrid factorisation ---------------------- */
/* Head: pivots order[0..nhead-1] eliminated in that order without pivoting
 * (multipliers stored in place, the pivot slot holds its reciprocal).  Tail:
 * the Schur complement on order[nhead..nb-1] is copied to `tail` and
 * inverted by gj_invert.  This is the factorisation of the workgroup kernel. */
static int hybrid_factor(int nb, double *a, const int *order, int nhead, double *tail, int *tpiv,
                         int f32) {
  if (f32)
    for (size_t i = 0; i < (size_t)nb * nb; i++) a[i] = (float)a[i];
  for (int h = 0; h < nhead; h++) {
    const int k = order[h];
    const double d = a[(size_t)k * nb + k];
    /* no pivoting in the head: a zero or non-finite pivot (in FP32, one whose
     * reciprocal is not a finite normal float) fails the setup */
    if (d == 0.0 || !isfinite(d)) return 1;
    double inv = 1.0 / d;
    if (f32) {
      inv = (float)inv;
      if (!isfinite(inv) || fabs(d) < FLT_MIN) return 1;
    }
    a[(size_t)k * nb + k] = inv;
    for (int hh = h + 1; hh < nb; hh++) {
      const int i = order[hh];
      double l = a[(size_t)i * nb + k] * inv;
      if (f32) l = (float)l;
      a[(size_t)i * nb + k] = l;
      if (l == 0.0) continue;
      for (int jj = h + 1; jj < nb; jj++) {
        const int j = order[jj];
        double v = a[(size_t)i * nb + j] - l * a[(size_t)k * nb + j];
        a[(size_t)i * nb + j] = f32 ? (float)v : v;
      }
    }
  }
  const int t = nb - nhead;
  for (int r = 0; r < t; r++)
    for (int c = 0; c < t; c++)
      tail[(size_t)r * t + c] = a[(size_t)order[nhead + r] * nb + order[nhead + c]];
  return gj_invert(t, tail, tpiv, f32);
}

static void hybrid_solve(int nb, const double *a, const int *order, int nhead, const double *tail,
                         const int *tpiv, double *b, double *work) {
  for (int h = 0; h < nhead; h++) {
    const int k = order[h];
    const double bk = b[k];
    for (int hh = h + 1; hh < nb; hh++) {
      const int i = order[hh];
      b[i] -= a[(size_t)i * nb + k] * bk;
    }
  }
  const int t = nb - nhead;
  double *bt = work;
  for (int r = 0; r < t; r++) bt[r] = b[order[nhead + r]];
  gj_apply(t, tail, tpiv, bt, work + t);
  for (int r = 0; r < t; r++) b[order[nhead + r]] = bt[r];
  for (int h = nhead - 1; h >= 0; h--) {
    const int k = order[h];
    double s = b[k];
    for (int jj = h + 1; jj < nb; jj++) {
      const int j = order[jj];
      s -= a[(size_t)k * nb + j] * b[j];
    }
    b[k] = s * a[(size_t)k * nb + k];
  }
}

/* ---------------- Jacobian: difference quotient or exact -------------- */
static int dq_jac(bdf_mem *m, const double *y, const double *fy, double *J) {
  const int n = m->n;
  double srur = sqrt(UROUND);
  double fnorm = wrms(n, fy, m->ewt);
  double minInc = (fnorm != 0.0) ? (MIN_INC_MULT * fabs(m->h) * UROUND * n * fnorm) : 1.0;
  double *yy = m->vtmp, *ft = m->tempv;
  memcpy(yy, y, sizeof(double) * n);
  for (int j = 0; j < n; j++) {
    double ys = yy[j];
    double inc = fmax(srur * fabs(ys), minInc / m->ewt[j]);
    yy[j] += inc;
    int ret = call_f(m, m->tn, yy, ft);
    m->nfeDQ++;
    if (ret != 0) return ret;
    yy[j] = ys;
    double ii = 1.0 / inc;
    for (int i = 0; i < n; i++) J[(size_t)i * n + j] = ii * ft[i] - ii * fy[i];
  }
  return 0;
}

/* linear-solver setup (dense): decide whether J is re-evaluated, form
 * M = I - gamma*J and factor it */
static int lsetup_hybrid(bdf_mem *m, int convfail, const double *ypred, int *jcur) {
  const int n = m->n, nb = m->nb;
  const size_t nn = (size_t)nb * nb;
  double dgamma = fabs(m->gamma / m->gammap - 1.0);
  int jbad = (m->nst == 0) || (m->nst > m->nstlj + MSBJ) ||
             ((convfail == CV_FAIL_BAD_J) && !(dgamma >= DGMAX_J)) || (convfail == CV_FAIL_OTHER);
  if (!jbad) {
    *jcur = 0;
  } else {
    m->nje++;
    m->nstlj = m->nst;
    *jcur = 1;
    if (m->jac(m->tn, ypred, m->savedJ, m->ud) != 0) return -1;
    if (m->mode & PDR_MODE_F32)
      for (size_t i = 0; i < nn; i++) m->savedJ[i] = (float)m->savedJ[i];
  }
  for (size_t i = 0; i < nn; i++) m->M[i] = -m->gamma * m->savedJ[i];
  for (int i = 0; i < n; i++) m->M[(size_t)i * nb + i] += 1.0;
  double xe = 0.0;
  for (int j = 0; j < n; j++) {
    m->M[(size_t)n * nb + j] = -m->bq[j];
    xe += m->bq[j] * ypred[j];
  }
  m->M[(size_t)n * nb + n] = 1.0;
  if (m->mode & PDR_MODE_SCALED) {
    for (int i = 0; i < n; i++) m->wsc[i] = m->ewt[i];
    m->wsc[n] = 1.0 / (m->rtol * fabs(xe) + m->atol);
    for (int i = 0; i < nb; i++)
      for (int j = 0; j < nb; j++) m->M[(size_t)i * nb + j] *= m->wsc[i] / m->wsc[j];
  }
  return hybrid_factor(nb, m->M, m->border_order, m->nhead, m->tail, m->tpiv,
                       (m->mode & PDR_MODE_F32) != 0) ? 1 : 0;
}

static int lsetup(bdf_mem *m, int convfail, const double *ypred, const double *fpred, int *jcur) {
  const int n = m->n;
  if (m->mode & PDR_MODE_HYBRID) return lsetup_hybrid(m, convfail, ypred, jcur);
  double dgamma = fabs(m->gamma / m->gammap - 1.0);
  int jbad = (m->nst == 0) || (m->nst > m->nstlj + MSBJ) ||
             ((convfail == CV_FAIL_BAD_J) && !(dgamma >= DGMAX_J)) || (convfail == CV_FAIL_OTHER);
  if (!jbad) {
    *jcur = 0;
    memcpy(m->M, m->savedJ, sizeof(double) * n * n);
  } else {
    m->nje++;
    m->nstlj = m->nst;
    *jcur = 1;
    int ret;
    if (!(m->mode & PDR_MODE_EXACT_JAC) || !m->jac) ret = dq_jac(m, ypred, fpred, m->M);
    else ret = m->jac(m->tn, ypred, m->M, m->ud);
    if (ret < 0) return -1;
    if (ret > 0) return 1;
    if (m->mode & PDR_MODE_F32)
      for (int i = 0; i < n * n; i++) m->M[i] = (float)m->M[i];
    memcpy(m->savedJ, m->M, sizeof(double) * n * n);
  }
  for (int i = 0; i < n * n; i++) m->M[i] *= -m->gamma;
  for (int i = 0; i < n; i++) m->M[(size_t)i * n + i] += 1.0;
  if (m->mode & PDR_MODE_SCALED) {
    /* factor W M W^-1 with W = diag(ewt): errors of the solve are then
     * measured in the same weighted norm as the Newton and error tests */
    for (int i = 0; i < n; i++) m->wsc[i] = m->ewt[i];
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) m->M[(size_t)i * n + j] *= m->wsc[i] / m->wsc[j];
  }
  if (m->mode & PDR_MODE_GJ) return gj_invert(n, m->M, m->piv, (m->mode & PDR_MODE_F32) != 0);
  if (m->mode & PDR_MODE_STATIC)
    return lu_factor_static(n, m->M, m->piv, (m->mode & PDR_MODE_F32) != 0) > 0 ? 1 : 0;
  int ier = lu_factor(n, m->M, m->piv, (m->mode & PDR_MODE_F32) != 0);
  return ier > 0 ? 1 : 0;
}

/* solve with the factored (possibly scaled) iteration matrix */
static void fsolve(bdf_mem *m, double *b) {
  const int n = m->n;
  if (m->mode & PDR_MODE_HYBRID) {
    const int sc = (m->mode & PDR_MODE_SCALED) != 0;
    double *bb = m->bwork;
    for (int i = 0; i < n; i++) bb[i] = sc ? b[i] * m->wsc[i] : b[i];
    bb[n] = 0.0;
    hybrid_solve(m->nb, m->M, m->border_order, m->nhead, m->tail, m->tpiv, bb, bb + m->nb);
    for (int i = 0; i < n; i++) b[i] = sc ? bb[i] / m->wsc[i] : bb[i];
    return;
  }
  if (m->mode & PDR_MODE_SCALED)
    for (int i = 0; i < n; i++) b[i] *= m->wsc[i];
  if (m->mode & PDR_MODE_GJ) gj_apply(n, m->M, m->piv, b, m->gjw);
  else if (m->mode & PDR_MODE_STATIC) lu_solve_static(n, m->M, m->piv, b);
  else lu_solve(n, m->M, m->piv, b);
  if (m->mode & PDR_MODE_SCALED)
    for (int i = 0; i < n; i++) b[i] /= m->wsc[i];
}

static void lsolve(bdf_mem *m, double *b) {
  const int n = m->n;
  int nref = (m->mode & PDR_MODE_HYBRID) ? 0 : (m->mode >> 2) & 3;
  double *rhs = m->vtmp;
  if (nref) memcpy(rhs, b, sizeof(double) * n);
  fsolve(m, b);
  /* iterative refinement against the unfactored matrix I - gammap*J
   * (the matrix the factorisation was formed with) */
  for (int it = 0; it < nref; it++) {
    for (int i = 0; i < n; i++) {
      const double *Ji = m->savedJ + (size_t)i * n;
      double acc = rhs[i] - b[i];
      for (int j = 0; j < n; j++) acc += m->gammap * Ji[j] * b[j];
      m->rsd[i] = acc;
    }
    fsolve(m, m->rsd);
    for (int i = 0; i < n; i++) b[i] += m->rsd[i];
  }
  if (m->gamrat != 1.0) {
    double s = 2.0 / (1.0 + m->gamrat);
    for (int i = 0; i < m->n; i++) b[i] *= s;
  }
}

/* ---------------- step machinery -------------------------------------- */

static void rescale(bdf_mem *m) {
  double factor = m->eta;
  for (int j = 1; j <= m->q; j++) {
    for (int i = 0; i < m->n; i++) m->zn[j][i] *= factor;
    factor *= m->eta;
  }
  m->h = m->hscale * m->eta;
  m->hscale = m->h;
  m->nscon = 0;
}

static void predict(bdf_mem *m) {
  m->tn += m->h;
  for (int k = 1; k <= m->q; k++)
    for (int j = m->q; j >= k; j--)
      for (int i = 0; i < m->n; i++) m->zn[j - 1][i] += m->zn[j][i];
}

static void restore(bdf_mem *m, double saved_t) {
  m->tn = saved_t;
  for (int k = 1; k <= m->q; k++)
    for (int j = m->q; j >= k; j--)
      for (int i = 0; i < m->n; i++) m->zn[j - 1][i] -= m->zn[j][i];
}

/* BDF coefficients l[] and error-test quantities tq[] for the current step
 * history (fixed-leading-coefficient form) */
static void set_bdf(bdf_mem *m) {
  const int q = m->q;
  double *l = m->l, *tau = m->tau, *tq = m->tq;
  double xi_inv = 1.0, xistar_inv = 1.0;
  l[0] = l[1] = 1.0;
  for (int i = 2; i <= q; i++) l[i] = 0.0;
  double alpha0 = -1.0, alpha0_hat = -1.0;
  double hsum = m->h;
  if (q > 1) {
    for (int j = 2; j < q; j++) {
      hsum += tau[j - 1];
      xi_inv = m->h / hsum;
      alpha0 -= 1.0 / j;
      for (int i = j; i >= 1; i--) l[i] += l[i - 1] * xi_inv;
    }
    alpha0 -= 1.0 / q;
    xistar_inv = -l[1] - alpha0;
    hsum += tau[q - 1];
    xi_inv = m->h / hsum;
    alpha0_hat = -l[1] - xi_inv;
    for (int i = q; i >= 1; i--) l[i] += l[i - 1] * xistar_inv;
  }
  /* error-test constants */
  double A1 = 1.0 - alpha0_hat + alpha0;
  double A2 = 1.0 + q * A1;
  tq[2] = fabs(A1 / (alpha0 * A2));
  tq[5] = fabs(A2 * xistar_inv / (l[q] * xi_inv));
  if (m->qwait == 1) {
    if (q > 1) {
      double C = xistar_inv / l[q];
      double A3 = alpha0 + 1.0 / q;
      double A4 = alpha0_hat + xi_inv;
      double Cpinv = (1.0 - A4 + A3) / A3;
      tq[1] = fabs(C * Cpinv);
    } else {
      tq[1] = 1.0;
    }
    hsum += tau[q];
    xi_inv = m->h / hsum;
    double A5 = alpha0 - 1.0 / (q + 1);
    double A6 = alpha0_hat - xi_inv;
    double Cppinv = (1.0 - A6 + A5) / A2;
    tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
  }
  tq[4] = NLSCOEF / tq[2];
}

static void set_coeffs(bdf_mem *m) {
  set_bdf(m);
  m->rl1 = 1.0 / m->l[1];
  m->gamma = m->h * m->rl1;
  if (m->nst == 0) m->gammap = m->gamma;
  m->gamrat = (m->nst > 0) ? m->gamma / m->gammap : 1.0;
}

static int newton_iter(bdf_mem *m) {
  const int n = m->n;
  int mm = 0;
  double del = 0.0, delp = 0.0;
  for (;;) {
    for (int i = 0; i < n; i++) {
      double tv = m->rl1 * m->zn[1][i] + m->acor[i];
      m->tempv[i] = m->gamma * m->ftemp[i] - tv;
    }
    lsolve(m, m->tempv);
    m->nni++;
    del = wrms(n, m->tempv, m->ewt);
    for (int i = 0; i < n; i++) {
      m->acor[i] += m->tempv[i];
      m->y[i] = m->zn[0][i] + m->acor[i];
    }
    if (mm > 0) m->crate = fmax(CRDOWN * m->crate, del / delp);
    double dcon = del * fmin(1.0, m->crate) / m->tq[4];
    if (dcon <= 1.0) {
      m->acnrm = (mm == 0) ? del : wrms(n, m->acor, m->ewt);
      m->jcur = 0;
      return 0;
    }
    mm++;
    if ((mm == NLS_MAXCOR) || ((mm >= 2) && (del > RDIV * delp))) {
      if (!m->jcur) return TRY_AGAIN;
      return CONV_FAIL;
    }
    delp = del;
    int ret = call_f(m, m->tn, m->y, m->ftemp);
    m->nfe++;
    if (ret < 0) return PDR_CV_RHSFUNC_FAIL;
    if (ret > 0) {
      if (!m->jcur) return TRY_AGAIN;
      return RHSFUNC_RECVR;
    }
  }
}

static int nls(bdf_mem *m, int nflag) {
  int convfail = ((nflag == FIRST_CALL) || (nflag == PREV_ERR_FAIL)) ? CV_NO_FAILURES : CV_FAIL_OTHER;
  int callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (m->nst == 0) ||
                  (m->nst >= m->nstlp + MSBP) || (fabs(m->gamrat - 1.0) > DGMAX);
  for (;;) {
    int ret = call_f(m, m->tn, m->zn[0], m->ftemp);
    m->nfe++;
    if (ret < 0) return PDR_CV_RHSFUNC_FAIL;
    if (ret > 0) return RHSFUNC_RECVR;
    if (callSetup) {
      int ier = lsetup(m, convfail, m->zn[0], m->ftemp, &m->jcur);
      m->nsetups++;
      callSetup = 0;
      m->gamrat = m->crate = 1.0;
      m->gammap = m->gamma;
      m->nstlp = m->nst;
      if (ier < 0) return PDR_CV_LSETUP_FAIL;
      if (ier > 0) return CONV_FAIL;
    }
    for (int i = 0; i < m->n; i++) { m->acor[i] = 0.0; m->y[i] = m->zn[0][i]; }
    int ier = newton_iter(m);
    if (ier != TRY_AGAIN) return ier;
    callSetup = 1;
    convfail = CV_FAIL_BAD_J;
  }
}

static int handle_nflag(bdf_mem *m, int *nflag, double saved_t, int *ncf) {
  int nf = *nflag;
  if (nf == 0) return DO_ERROR_TEST;
  m->ncfn++;
  restore(m, saved_t);
  if (nf == PDR_CV_LSETUP_FAIL || nf == PDR_CV_RHSFUNC_FAIL) return nf;
  (*ncf)++;
  m->etamax = 1.0;
  if ((fabs(m->h) <= m->hmin * ONEPSM) || (*ncf == MXNCF)) {
    if (nf == CONV_FAIL) return PDR_CV_CONV_FAILURE;
    return PDR_CV_REPTD_RHSFUNC_ERR;
  }
  m->eta = fmax(ETACF, m->hmin / fabs(m->h));
  *nflag = PREV_CONV_FAIL;
  rescale(m);
  return PREDICT_AGAIN;
}

/* order change of the Nordsieck array (BDF) */
static void adjust_order(bdf_mem *m, int deltaq) {
  const int n = m->n, q = m->q;
  double *l = m->l;
  if ((q == 2) && (deltaq != 1)) return;
  if (deltaq == 1) {
    for (int i = 0; i <= m->qmax; i++) l[i] = 0.0;
    l[2] = 1.0;
    double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0;
    double hsum = m->hscale;
    if (q > 1) {
      for (int j = 1; j < q; j++) {
        hsum += m->tau[j + 1];
        double xi = hsum / m->hscale;
        prod *= xi;
        alpha0 -= 1.0 / (j + 1);
        alpha1 += 1.0 / xi;
        for (int i = j + 2; i >= 2; i--) l[i] = l[i] * xiold + l[i - 1];
        xiold = xi;
      }
    }
    double A1 = (-alpha0 - alpha1) / prod;
    double *zl = m->zn[m->L];
    const double *sa = m->zn[m->indx_acor];
    /* zn[L] = A1 * saved acor (zn[L] and zn[indx_acor] may alias at q = qmax-1) */
    for (int i = 0; i < n; i++) zl[i] = A1 * sa[i];
    for (int j = 2; j <= q; j++)
      for (int i = 0; i < n; i++) m->zn[j][i] += l[j] * zl[i];
  } else {
    for (int i = 0; i <= m->qmax; i++) l[i] = 0.0;
    l[2] = 1.0;
    double hsum = 0.0;
    for (int j = 1; j <= q - 2; j++) {
      hsum += m->tau[j];
      double xi = hsum / m->hscale;
      for (int i = j + 2; i >= 2; i--) l[i] = l[i] * xi + l[i - 1];
    }
    for (int j = 2; j < q; j++)
      for (int i = 0; i < n; i++) m->zn[j][i] -= l[j] * m->zn[q][i];
  }
}

static void adjust_params(bdf_mem *m) {
  if (m->qprime != m->q) {
    adjust_order(m, m->qprime - m->q);
    m->q = m->qprime;
    m->L = m->q + 1;
    m->qwait = m->L;
  }
  rescale(m);
}

static int do_error_test(bdf_mem *m, int *nflag, double saved_t, int *nef, double *dsmp) {
  double dsm = m->acnrm * m->tq[2];
  *dsmp = dsm;
  if (dsm <= 1.0) return 0;
  (*nef)++;
  m->netf++;
  *nflag = PREV_ERR_FAIL;
  restore(m, saved_t);
  if ((fabs(m->h) <= m->hmin * ONEPSM) || (*nef == MXNEF)) return PDR_CV_ERR_FAILURE;
  m->etamax = 1.0;
  if (*nef <= MXNEF1) {
    m->eta = 1.0 / (pow(BIAS2 * dsm, 1.0 / m->L) + ADDON);
    m->eta = fmax(ETAMIN, fmax(m->eta, m->hmin / fabs(m->h)));
    if (*nef >= SMALL_NEF) m->eta = fmin(m->eta, ETAMXF);
    rescale(m);
    return TRY_AGAIN;
  }
  if (m->q > 1) {
    m->eta = fmax(ETAMIN, m->hmin / fabs(m->h));
    adjust_order(m, -1);
    m->L = m->q;
    m->q--;
    m->qwait = m->L;
    rescale(m);
    return TRY_AGAIN;
  }
  m->eta = fmax(ETAMIN, m->hmin / fabs(m->h));
  m->h *= m->eta;
  m->hscale = m->h;
  m->qwait = LONG_WAIT;
  m->nscon = 0;
  int ret = call_f(m, m->tn, m->zn[0], m->tempv);
  m->nfe++;
  if (ret < 0) return PDR_CV_RHSFUNC_FAIL;
  if (ret > 0) return PDR_CV_UNREC_RHSFUNC_ERR;
  for (int i = 0; i < m->n; i++) m->zn[1][i] = m->h * m->tempv[i];
  return TRY_AGAIN;
}

static void complete_step(bdf_mem *m) {
  m->nst++;
  m->nscon++;
  m->hu = m->h;
  m->qu = m->q;
  for (int i = m->q; i >= 2; i--) m->tau[i] = m->tau[i - 1];
  if ((m->q == 1) && (m->nst > 1)) m->tau[2] = m->tau[1];
  m->tau[1] = m->h;
  for (int j = 0; j <= m->q; j++)
    for (int i = 0; i < m->n; i++) m->zn[j][i] += m->l[j] * m->acor[i];
  m->qwait--;
  if ((m->qwait == 1) && (m->q != m->qmax)) {
    memcpy(m->zn[m->qmax], m->acor, sizeof(double) * m->n);
    m->saved_tq5 = m->tq[5];
    m->indx_acor = m->qmax;
  }
}

static void set_eta(bdf_mem *m) {
  if (m->eta < THRESH) {
    m->eta = 1.0;
    m->hprime = m->h;
  } else {
    m->eta = fmin(m->eta, m->etamax);
    m->eta /= fmax(1.0, fabs(m->h) * m->hmax_inv * m->eta);
    m->hprime = m->h * m->eta;
    if (m->qprime < m->q) m->nscon = 0;
  }
}

static void prepare_next_step(bdf_mem *m, double dsm) {
  if (m->etamax == 1.0) {
    m->qwait = (m->qwait > 2) ? m->qwait : 2;
    m->qprime = m->q;
    m->hprime = m->h;
    m->eta = 1.0;
    return;
  }
  double etaq = 1.0 / (pow(BIAS2 * dsm, 1.0 / m->L) + ADDON);
  if (m->qwait != 0) {
    m->eta = etaq;
    m->qprime = m->q;
    set_eta(m);
    return;
  }
  m->qwait = 2;
  double etaqm1 = 0.0, etaqp1 = 0.0;
  if (m->q > 1) {
    double ddn = wrms(m->n, m->zn[m->q], m->ewt) * m->tq[1];
    etaqm1 = 1.0 / (pow(BIAS1 * ddn, 1.0 / m->q) + ADDON);
  }
  if (m->q != m->qmax && m->saved_tq5 != 0.0) {
    double cquot = (m->tq[5] / m->saved_tq5) * pow(m->h / m->tau[2], (double)m->L);
    for (int i = 0; i < m->n; i++) m->tempv[i] = m->acor[i] - cquot * m->zn[m->qmax][i];
    double dup = wrms(m->n, m->tempv, m->ewt) * m->tq[3];
    etaqp1 = 1.0 / (pow(BIAS3 * dup, 1.0 / (m->L + 1)) + ADDON);
  }
  /* choose the largest of the three candidate ratios */
  double etam = fmax(etaqm1, fmax(etaq, etaqp1));
  if (etam < THRESH) {
    m->eta = 1.0;
    m->qprime = m->q;
  } else if (etam == etaq) {
    m->eta = etaq;
    m->qprime = m->q;
  } else if (etam == etaqm1) {
    m->eta = etaqm1;
    m->qprime = m->q - 1;
  } else {
    m->eta = etaqp1;
    m->qprime = m->q + 1;
    memcpy(m->zn[m->qmax], m->acor, sizeof(double) * m->n);
  }
  set_eta(m);
}

static int bdf_step(bdf_mem *m) {
  double saved_t = m->tn, dsm = 0.0;
  int ncf = 0, nef = 0, nflag = FIRST_CALL;
  if ((m->nst > 0) && (m->hprime != m->h)) adjust_params(m);
  for (;;) {
    predict(m);
    set_coeffs(m);
    nflag = nls(m, nflag);
    int kflag = handle_nflag(m, &nflag, saved_t, &ncf);
    if (kflag == PREDICT_AGAIN) continue;
    if (kflag != DO_ERROR_TEST) return kflag;
    int eflag = do_error_test(m, &nflag, saved_t, &nef, &dsm);
    if (eflag == TRY_AGAIN) continue;
    if (eflag != 0) return eflag;
    break;
  }
  complete_step(m);
  prepare_next_step(m, dsm);
  m->etamax = (m->nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
  for (int i = 0; i < m->n; i++) m->acor[i] *= m->tq[2];
  return 0;
}

/* initial step size estimate */
static int hin(bdf_mem *m, double tout) {
  const int n = m->n;
  double tdiff = tout - m->tn;
  if (tdiff == 0.0) return PDR_CV_TOO_CLOSE;
  int sign = (tdiff > 0.0) ? 1 : -1;
  double tdist = fabs(tdiff);
  double tround = UROUND * fmax(fabs(m->tn), fabs(tout));
  if (tdist < 2.0 * tround) return PDR_CV_TOO_CLOSE;
  double hlb = HLB_FACTOR * tround;
  /* upper bound: forward Euler may change no component by more than
   * HUB_FACTOR of its size (plus its absolute tolerance) */
  double hub_inv = 0.0;
  for (int i = 0; i < n; i++) {
    double den = HUB_FACTOR * fabs(m->zn[0][i]) + 1.0 / m->ewt[i];
    double v = fabs(m->zn[1][i]) / den;
    if (v > hub_inv) hub_inv = v;
  }
  double hub = HUB_FACTOR * tdist;
  if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
  double hg = sqrt(hlb * hub);
  if (hub < hlb) {
    m->h = (sign == -1) ? -hg : hg;
    return 0;
  }
  double hnew = hg;
  for (int count1 = 1; count1 <= MAX_ITERS; count1++) {
    int hgOK = 0;
    double yddnrm = 0.0;
    for (int count2 = 1; count2 <= MAX_ITERS; count2++) {
      double hgs = hg * sign;
      for (int i = 0; i < n; i++) m->y[i] = m->zn[0][i] + hgs * m->zn[1][i];
      int ret = call_f(m, m->tn + hgs, m->y, m->tempv);
      m->nfe++;
      if (ret < 0) return PDR_CV_RHSFUNC_FAIL;
      if (ret == 0) {
        for (int i = 0; i < n; i++) m->tempv[i] = (m->tempv[i] - m->zn[1][i]) * (1.0 / hgs);
        yddnrm = wrms(n, m->tempv, m->ewt);
        hgOK = 1;
        break;
      }
      hg *= 0.2;
    }
    if (!hgOK) return PDR_CV_REPTD_RHSFUNC_ERR;
    hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
    if (count1 == MAX_ITERS) break;
    double hrat = hnew / hg;
    if ((hrat > 0.5) && (hrat < 2.0)) break;
    if ((count1 > 1) && (hrat > 2.0)) { hnew = hg; break; }
    hg = hnew;
  }
  double h0 = H_BIAS * hnew;
  if (h0 < hlb) h0 = hlb;
  if (h0 > hub) h0 = hub;
  m->h = (sign == -1) ? -h0 : h0;
  return 0;
}

int bdf_get_dky(bdf_mem *m, double t, double *dky) {
  double tfuzz = FUZZ_FACTOR * UROUND * (fabs(m->tn) + fabs(m->hu));
  if (m->hu < 0.0) tfuzz = -tfuzz;
  double tp = m->tn - m->hu - tfuzz;
  double tn1 = m->tn + tfuzz;
  if ((t - tp) * (t - tn1) > 0.0) return PDR_CV_BAD_T;
  double s = (t - m->tn) / m->h;
  for (int i = 0; i < m->n; i++) dky[i] = m->zn[m->q][i];
  for (int j = m->q - 1; j >= 0; j--)
    for (int i = 0; i < m->n; i++) dky[i] = m->zn[j][i] + s * dky[i];
  return 0;
}

/* One CVode(tout, CV_NORMAL) call.  On success *tret = tout and yout holds
 * the interpolated solution; on failure *tret = tn and yout = zn[0]. */
int bdf_solve(bdf_mem *m, double tout, double *yout, double *tret) {
  const int n = m->n;
  if (m->first) {
    set_ewt(m, m->zn[0]);
    int ret = call_f(m, m->tn, m->zn[0], m->zn[1]);
    m->nfe++;
    if (ret != 0) { *tret = m->tn; return PDR_CV_FIRST_RHSFUNC_ERR; }
    int hf = hin(m, tout);
    if (hf != 0) { *tret = m->tn; return hf; }
    double rh = fabs(m->h) * m->hmax_inv;
    if (rh > 1.0) m->h /= rh;
    if (fabs(m->h) < m->hmin) m->h *= m->hmin / fabs(m->h);
    m->hscale = m->h;
    m->hprime = m->h;
    for (int i = 0; i < n; i++) m->zn[1][i] *= m->h;
    m->first = 0;
  } else if ((m->tn - tout) * m->h >= 0.0) {
    /* tout already passed by an earlier step: interpolate only */
    *tret = tout;
    return bdf_get_dky(m, tout, yout);
  }
  long nstloc = 0;
  for (;;) {
    if (m->nst > 0) set_ewt(m, m->zn[0]);
    if ((m->mxstep > 0) && (nstloc >= m->mxstep)) {
      *tret = m->tn;
      memcpy(yout, m->zn[0], sizeof(double) * n);
      return PDR_CV_TOO_MUCH_WORK;
    }
    double nrm = wrms(n, m->zn[0], m->ewt);
    if (UROUND * nrm > 1.0) {
      *tret = m->tn;
      memcpy(yout, m->zn[0], sizeof(double) * n);
      return PDR_CV_TOO_MUCH_ACC;
    }
    int kflag = bdf_step(m);
    if (kflag != 0) {
      *tret = m->tn;
      memcpy(yout, m->zn[0], sizeof(double) * n);
      return kflag;
    }
    nstloc++;
    if ((m->tn - tout) * m->h >= 0.0) {
      *tret = tout;
      bdf_get_dky(m, tout, yout);
      return 0;
    }
  }
}

/* The bordered factorisation's status for one matrix (0 = factorised, 1 =
 * setup failure), for tests of its pivot rule; a[nb*nb] is overwritten. */
int pdr_test_hybrid_factor(int nb, double *a, const int *order, int nhead, int f32) {
  const int t = nb - nhead;
  double *tail = malloc(sizeof(double) * (size_t)(t > 0 ? t * t : 1));
  int *tpiv = malloc(sizeof(int) * (size_t)(t > 0 ? t : 1));
  int rc = (tail && tpiv) ? hybrid_factor(nb, a, order, nhead, tail, tpiv, f32) : -1;
  free(tail);
  free(tpiv);
  return rc;
}
