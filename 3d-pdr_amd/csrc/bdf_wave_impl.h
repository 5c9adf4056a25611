// Batched variable-order BDF, one stiff system per wavefront (wave64).
//
// Included once per network inside a namespace that defines `namespace NET`
// (a generated csrc/gen/net_<name>.h).  Replaces, for one grid point per
// wavefront, the reference's per-point CVODE solve (calculate_abundances.c:
// 132-326) and the generated RHS (odes_*.c:26-148).
//
// Integrator: CVODE 2.7.0's BDF (Nordsieck array, fixed-leading-coefficient
// orders 1-5, local error test, step/order selection, Newton iteration with
// the same convergence/divergence tests and setup/Jacobian refresh rules,
// h0 estimate, CV_NORMAL output by interpolation) -- the same restatement
// the CPU oracle carries, run with the exact Jacobian.  The linear solve
// differs from CVODE's dense LU: the iteration matrix is scaled by the error
// weights (W M W^-1), inverted by Gauss-Jordan with partial pivoting in
// registers, and every solve is followed by three refinement steps with an
// FP64 residual of I - gammap*J.  With atol = 1e-30 the plain solve's
// rounding error in trace species stalls the Newton iteration; the scaled,
// refined solve removes those spurious failures without changing the
// converged solution.  MT (the matrix type of J and of the inverse) is
// double for the plain FP64 form and float for the mixed form; the residual,
// Newton update, error test and all n-vectors stay FP64 in both.
//
// Layout, lane L of the wave: row i = L & 31, half hh = L >> 5.  The lane
// holds columns hh*16 .. hh*16+15 of row i of J and of the inverse; n-vectors
// are held as the lane's row value (duplicated in both halves).  Step-control
// scalars are wave-uniform and kept in the wave's LDS record.

#include "wave_ops.h"

using namespace pdrw;

constexpr int QMAX = 5;
constexpr double UROUND = 2.220446049250313e-16;
constexpr double ETAMX1 = 10000.0, ETAMX2 = 10.0, ETAMX3 = 10.0, ETAMXF = 0.2, ETAMIN = 0.1;
constexpr double ETACF = 0.25, ADDON = 0.000001, BIAS1 = 6.0, BIAS2 = 6.0, BIAS3 = 10.0;
constexpr double THRESH = 1.5, CRDOWN = 0.3, DGMAX = 0.3, RDIV = 2.0;
constexpr double DGMAX_J = 0.2, NLSCOEF = 0.1, HLB_FACTOR = 100.0, HUB_FACTOR = 0.1;
constexpr double H_BIAS = 0.5;
constexpr int SMALL_NST = 10, MXNCF = 10, MXNEF = 7, MXNEF1 = 3, SMALL_NEF = 2, LONG_WAIT = 10;
constexpr int NLS_MAXCOR = 3, MSBP = 20, MSBJ = 50, MAX_ITERS = 4, NREFINE = 3;
constexpr int WAVES = 4;  // wavefronts per block

enum { CONV_FAIL = 4, TRY_AGAIN = 5, FIRST_CALL = 6, PREV_CONV_FAIL = 7, PREV_ERR_FAIL = 8 };
enum { CV_NO_FAILURES = 0, CV_FAIL_BAD_J = 1, CV_FAIL_OTHER = 2 };

// uniform integrator state of one wavefront (lives in LDS)
struct BdfU {
  double tn, h, hscale, hprime, hu, eta, etamax;
  double tau[QMAX + 2], tq[6], l[QMAX + 1];
  double rl1, gamma, gammap, gamrat, crate, acnrm, saved_tq5, xe_last;
  int q, qprime, L, qwait, nscon, jcur;
  int nst, nfe, nsetups, nje, nni, ncfn, netf, nstlp, nstlj;
};

// per-wavefront shared memory
struct WaveLDS {
  double X[64];                    // factor vector: y, x_e, 1
  double F[2 * NET::NSLOT + 1];    // fluxes (RHS) or flux partials (Jacobian); last = 0
  double P[65];                    // RHS lane partials; P[64] = 0
  double V[64];                    // vector exchange
  double V2[64];
  double C[32];                    // pivot-column exchange
  BdfU u;
};

// per-block shared memory: the RHS term table is the same for every wave
struct BlockLDS {
  uint32_t rterm[NET::TM * 64];
  WaveLDS w[WAVES];
};

struct Lane {
  int L, i, hh;
  bool row_ok;      // i < NEQ
  bool owner;       // hh == 0 && row_ok: the lane that publishes row values
};

struct Tables {
  uint32_t fab[NET::ROUNDS];
  int rbase, rnl;
  double q;          // charge of row i
  const uint32_t *rterm;  // block LDS copy
};

template <typename MT>
struct Ctx {
  Lane ln;
  Tables tb;
  double Kp[NET::ROUNDS];
  WaveLDS *S;
  double rtol, atol;
};

// per-lane integrator vectors and matrices (registers)
template <typename MT>
struct Bdf {
  double zn[QMAX + 1];
  double ewt, acor, y, ftemp, wsc;
  MT J[16], A[16];
  int invp;
};

// ---------------------------------------------------------------------
// RHS: returns ydot_i for this lane's row; x_e (uniform) via xe_out
// ---------------------------------------------------------------------
template <typename MT>
__device__ __forceinline__ double rhs(Ctx<MT> &c, double y, double &xe_out) {
  const Lane &ln = c.ln;
  WaveLDS &S = *c.S;
  if (ln.owner) S.X[ln.i] = y;
  double xe = wave_sum(ln.owner ? c.tb.q * y : 0.0);
  xe_out = xe;
  if (ln.L == 0) S.X[NET::NEQ] = xe;
  wsync();
#pragma unroll
  for (int r = 0; r < NET::ROUNDS; r++) {
    uint32_t f = c.tb.fab[r];
    S.F[r * 64 + ln.L] = c.Kp[r] * S.X[f & 0xff] * S.X[(f >> 8) & 0xff];
  }
  wsync();
  double acc = 0.0;
#pragma unroll
  for (int t = 0; t < NET::TM; t++) {
    uint32_t w = c.tb.rterm[t * 64 + ln.L];
    acc = fma((double)(int16_t)(w >> 16), S.F[w & 0xffff], acc);
  }
  S.P[ln.L] = acc;
  wsync();
  double yd = 0.0;
#pragma unroll
  for (int u = 0; u < NET::PM; u++) yd += (u < c.tb.rnl) ? S.P[min(c.tb.rbase + u, 64)] : 0.0;
  wsync();
  return yd;
}

// ---------------------------------------------------------------------
// exact Jacobian at the factor vector currently in S.X: row i, columns
// hh*16.., including the x_e chain term g_i * charge_j
// ---------------------------------------------------------------------
template <typename MT>
__device__ __forceinline__ void jacobian(Ctx<MT> &c, MT (&Jl)[16]) {
  const Lane &ln = c.ln;
  WaveLDS &S = *c.S;
#pragma unroll
  for (int r = 0; r < NET::ROUNDS; r++) {
    uint32_t f = c.tb.fab[r];
    int s = r * 64 + ln.L;
    S.F[2 * s] = c.Kp[r] * S.X[(f >> 8) & 0xff];   // d flux / d X[a]
    S.F[2 * s + 1] = c.Kp[r] * S.X[f & 0xff];      // d flux / d X[b]
  }
  wsync();
  double acc[17];
  int off = 0;
#pragma unroll
  for (int j = 0; j < 17; j++) {
    double a = 0.0;
#pragma nounroll
    for (int t = 0; t < NET::JT[j]; t++) {
      uint32_t w = NET::jac_term[(off + t) * 64 + ln.L];
      a = fma((double)(int16_t)(w >> 16), S.F[w & 0xffff], a);
    }
    off += NET::JT[j];
    acc[j] = a;
  }
#pragma unroll
  for (int j = 0; j < 16; j++) Jl[j] = (MT)fma(acc[16], (double)NET::charge[ln.hh * 16 + j], acc[j]);
  wsync();
}

// ---------------------------------------------------------------------
// Gauss-Jordan inverse (in place, partial pivoting) of the scaled matrix.
// On return A holds the inverse in pivot order; invp = step at which this
// row was the pivot.  Returns false if singular.
// ---------------------------------------------------------------------
template <typename MT>
__device__ __forceinline__ bool gj_invert(const Lane &ln, WaveLDS &S, MT (&A)[16], int &invp) {
  bool used = !ln.row_ok;
  invp = ln.i;
  MT *Vm = reinterpret_cast<MT *>(S.V);
  MT *Cm = reinterpret_cast<MT *>(S.C);
#pragma unroll
  for (int k = 0; k < NET::NEQ; k++) {
    const int hk = k >> 4, jk = k & 15;
    double v = (ln.hh == hk && !used) ? fabs((double)A[jk]) : -1.0;
    double vmax;
    int p;
    wave_argmax(v, ln.i, vmax, p);
    if (!(vmax > 0.0)) return false;
    if (ln.hh == hk) Cm[ln.i] = A[jk];
    if (ln.i == p) {
#pragma unroll
      for (int j = 0; j < 16; j++) Vm[ln.hh * 16 + j] = A[j];
    }
    wsync();
    const MT fi = Cm[ln.i];
    const MT d = (MT)1.0 / Cm[p];
    const bool isp = (ln.i == p);
    if (isp) { used = true; invp = k; }
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const MT pj = Vm[ln.hh * 16 + j] * d;
      const bool isk = (ln.hh == hk) && (j == jk);
      MT nv = isp ? pj : fma(-fi, pj, A[j]);
      MT kv = isp ? d : -fi * d;
      A[j] = isk ? kv : nv;
    }
    wsync();
  }
  return true;
}

// y = (I - g*J) b for row values b (the unscaled iteration matrix)
template <typename MT>
__device__ __forceinline__ double imj_apply(const Lane &ln, WaveLDS &S, const MT (&J)[16], double g, double b) {
  if (ln.hh == 0) S.V[ln.i] = ln.row_ok ? b : 0.0;
  wsync();
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < 16; j++) acc = fma((double)J[j], S.V[ln.hh * 16 + j], acc);
  acc += shfl_xor_d(acc, 32);
  wsync();
  return ln.row_ok ? b - g * acc : 0.0;
}

// x = M^-1 b with the pivot-ordered in-place inverse
template <typename MT>
__device__ __forceinline__ double inv_apply(const Lane &ln, WaveLDS &S, const MT (&A)[16], int invp, double b) {
  if (ln.hh == 0) S.V[invp] = ln.row_ok ? b : 0.0;
  wsync();
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < 16; j++) acc = fma((double)A[j], S.V[ln.hh * 16 + j], acc);
  acc += shfl_xor_d(acc, 32);
  if (ln.hh == 0) S.V2[invp] = acc;
  wsync();
  double x = S.V2[ln.i];
  wsync();
  return ln.row_ok ? x : 0.0;
}

// ---------------------------------------------------------------------
// integrator pieces
// ---------------------------------------------------------------------
template <typename MT>
__device__ __forceinline__ double wrms(const Ctx<MT> &c, double v, double w) {
  double p = c.ln.owner ? v * w : 0.0;
  return sqrt(wave_sum(p * p) / NET::NEQ);
}

template <typename MT>
__device__ __forceinline__ double f_eval(Ctx<MT> &c, double y) {
  double xe;
  double yd = rhs(c, y, xe);
  if (c.ln.L == 0) c.S->u.xe_last = xe;
  return yd;
}

template <int NA>
__device__ __forceinline__ double selz(const double (&a)[NA], int k) {
  double v = a[0];
#pragma unroll
  for (int j = 1; j < NA; j++) v = (k == j) ? a[j] : v;
  return v;
}

template <typename MT>
__device__ __forceinline__ void rescale(BdfU &u, Bdf<MT> &b) {
  const double eta = u.eta;
  const int q = u.q;
  double factor = eta;
#pragma unroll
  for (int j = 1; j <= QMAX; j++) {
    if (j <= q) {
      b.zn[j] *= factor;
      factor *= eta;
    }
  }
  u.h = u.hscale * eta;
  u.hscale = u.h;
  u.nscon = 0;
}

template <typename MT>
__device__ __forceinline__ void predict(BdfU &u, Bdf<MT> &b) {
  u.tn += u.h;
  const int q = u.q;
#pragma unroll
  for (int k = 1; k <= QMAX; k++)
#pragma unroll
    for (int j = QMAX; j >= k; j--)
      if (j <= q && k <= q) b.zn[j - 1] += b.zn[j];
}

template <typename MT>
__device__ __forceinline__ void restore(BdfU &u, Bdf<MT> &b, double saved_t) {
  u.tn = saved_t;
  const int q = u.q;
#pragma unroll
  for (int k = 1; k <= QMAX; k++)
#pragma unroll
    for (int j = QMAX; j >= k; j--)
      if (j <= q && k <= q) b.zn[j - 1] -= b.zn[j];
}

// BDF coefficients l[] and error-test constants tq[] (uniform scalar code)
__device__ void set_coeffs(BdfU &u) {
  const int q = u.q;
  const double h = u.h;
  double l[QMAX + 1], tau[QMAX + 2];
#pragma unroll
  for (int i = 0; i <= QMAX + 1; i++) tau[i] = u.tau[i];
  double xi_inv = 1.0, xistar_inv = 1.0;
  l[0] = l[1] = 1.0;
#pragma unroll
  for (int i = 2; i <= QMAX; i++) l[i] = 0.0;
  double alpha0 = -1.0, alpha0_hat = -1.0;
  double hsum = h;
  if (q > 1) {
#pragma unroll
    for (int j = 2; j < QMAX; j++) {
      if (j < q) {
        hsum += tau[j - 1];
        xi_inv = h / hsum;
        alpha0 -= 1.0 / j;
#pragma unroll
        for (int i = QMAX; i >= 1; i--)
          if (i <= j) l[i] += l[i - 1] * xi_inv;
      }
    }
    alpha0 -= 1.0 / q;
    xistar_inv = -l[1] - alpha0;
    hsum += selz(tau, q - 1);
    xi_inv = h / hsum;
    alpha0_hat = -l[1] - xi_inv;
#pragma unroll
    for (int i = QMAX; i >= 1; i--)
      if (i <= q) l[i] += l[i - 1] * xistar_inv;
  }
  const double lq = selz(l, q);
  double A1 = 1.0 - alpha0_hat + alpha0;
  double A2 = 1.0 + q * A1;
  double tq2 = fabs(A1 / (alpha0 * A2));
  u.tq[2] = tq2;
  u.tq[5] = fabs(A2 * xistar_inv / (lq * xi_inv));
  if (u.qwait == 1) {
    if (q > 1) {
      double C = xistar_inv / lq;
      double A3 = alpha0 + 1.0 / q;
      double A4 = alpha0_hat + xi_inv;
      double Cpinv = (1.0 - A4 + A3) / A3;
      u.tq[1] = fabs(C * Cpinv);
    } else {
      u.tq[1] = 1.0;
    }
    hsum += selz(tau, q);
    xi_inv = h / hsum;
    double A5 = alpha0 - 1.0 / (q + 1);
    double A6 = alpha0_hat - xi_inv;
    double Cppinv = (1.0 - A6 + A5) / A2;
    u.tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
  }
  u.tq[4] = NLSCOEF / tq2;
#pragma unroll
  for (int i = 0; i <= QMAX; i++) u.l[i] = l[i];
  u.rl1 = 1.0 / l[1];
  u.gamma = h * u.rl1;
  if (u.nst == 0) u.gammap = u.gamma;
  u.gamrat = (u.nst > 0) ? u.gamma / u.gammap : 1.0;
}

// linear-solver setup: Jacobian refresh decision, scaled Gauss-Jordan
template <typename MT>
__device__ int lsetup(Ctx<MT> &c, BdfU &u, Bdf<MT> &b, int convfail) {
  const Lane &ln = c.ln;
  WaveLDS &S = *c.S;
  double dgamma = fabs(u.gamma / u.gammap - 1.0);
  bool jbad = (u.nst == 0) || (u.nst > u.nstlj + MSBJ) ||
              ((convfail == CV_FAIL_BAD_J) && (dgamma < DGMAX_J)) || (convfail == CV_FAIL_OTHER);
  if (jbad) {
    const int nst = u.nst;
    if (ln.L == 0) {
      u.nje++;
      u.nstlj = nst;
      u.jcur = 1;
    }
    // factor vector at the predicted state (zn[0])
    if (ln.owner) S.X[ln.i] = b.zn[0];
    double xe = wave_sum(ln.owner ? c.tb.q * b.zn[0] : 0.0);
    if (ln.L == 0) S.X[NET::NEQ] = xe;
    wsync();
    jacobian(c, b.J);
  } else {
    if (ln.L == 0) u.jcur = 0;
  }
  const double gamma = u.gamma;
  // M = I - gamma J, scaled: W M W^-1 (w = error weights)
  b.wsc = ln.row_ok ? b.ewt : 1.0;
  if (ln.hh == 0) S.V[ln.i] = b.wsc;
  wsync();
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int col = ln.hh * 16 + j;
    const double wj = S.V[col];
    double m = -gamma * (double)b.J[j];
    if (col == ln.i) m += 1.0;
    if (!ln.row_ok) m = (col == ln.i) ? 1.0 : 0.0;
    b.A[j] = (MT)(m * (b.wsc / wj));
  }
  wsync();
  bool ok = gj_invert(ln, S, b.A, b.invp);
  return ok ? 0 : 1;
}

// solve M delta = r: scaled inverse, NREFINE refinement steps against
// I - gammap*J (the matrix the inverse was formed from), then CVODE's
// 2/(1+gamrat) correction for a changed gamma
template <typename MT>
__device__ double lsolve(Ctx<MT> &c, const BdfU &u, const Bdf<MT> &b, double r) {
  const double gammap = u.gammap, gamrat = u.gamrat;
  double x = inv_apply(c.ln, *c.S, b.A, b.invp, r * b.wsc) / b.wsc;
#pragma unroll
  for (int it = 0; it < NREFINE; it++) {
    double res = r - imj_apply(c.ln, *c.S, b.J, gammap, x);
    x += inv_apply(c.ln, *c.S, b.A, b.invp, res * b.wsc) / b.wsc;
  }
  if (gamrat != 1.0) x *= 2.0 / (1.0 + gamrat);
  return c.ln.row_ok ? x : 0.0;
}

template <typename MT>
__device__ int newton_iter(Ctx<MT> &c, BdfU &u, Bdf<MT> &b) {
  int m = 0;
  double del = 0.0, delp = 0.0;
  for (;;) {
    double rhs_v = u.gamma * b.ftemp - (u.rl1 * b.zn[1] + b.acor);
    double dx = lsolve(c, u, b, rhs_v);
    del = wrms(c, dx, b.ewt);
    b.acor += dx;
    b.y = b.zn[0] + b.acor;
    double crate = u.crate;
    if (m > 0) crate = fmax(CRDOWN * crate, del / delp);
    const double dcon = del * fmin(1.0, crate) / u.tq[4];
    const double acn = wrms(c, b.acor, b.ewt);
    wsync();
    if (c.ln.L == 0) {
      u.nni++;
      u.crate = crate;
    }
    if (dcon <= 1.0) {
      if (c.ln.L == 0) {
        u.acnrm = (m == 0) ? del : acn;
        u.jcur = 0;
      }
      wsync();
      return 0;
    }
    m++;
    if ((m == NLS_MAXCOR) || ((m >= 2) && (del > RDIV * delp))) {
      int jc = u.jcur;
      wsync();
      return jc ? CONV_FAIL : TRY_AGAIN;
    }
    delp = del;
    b.ftemp = f_eval(c, b.y);
    if (c.ln.L == 0) u.nfe++;
    wsync();
  }
}

template <typename MT>
__device__ int nls(Ctx<MT> &c, BdfU &u, Bdf<MT> &b, int nflag) {
  int convfail = ((nflag == FIRST_CALL) || (nflag == PREV_ERR_FAIL)) ? CV_NO_FAILURES : CV_FAIL_OTHER;
  bool callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (u.nst == 0) ||
                   (u.nst >= u.nstlp + MSBP) || (fabs(u.gamrat - 1.0) > DGMAX);
  for (;;) {
    b.ftemp = f_eval(c, b.zn[0]);
    if (c.ln.L == 0) u.nfe++;
    wsync();
    if (callSetup) {
      int ier = lsetup(c, u, b, convfail);
      wsync();
      if (c.ln.L == 0) {
        u.nsetups++;
        u.gamrat = 1.0;
        u.crate = 1.0;
        u.gammap = u.gamma;
        u.nstlp = u.nst;
      }
      wsync();
      callSetup = false;
      if (ier > 0) return CONV_FAIL;
    }
    b.acor = 0.0;
    b.y = b.zn[0];
    int ier = newton_iter(c, u, b);
    if (ier != TRY_AGAIN) return ier;
    callSetup = true;
    convfail = CV_FAIL_BAD_J;
  }
}

template <typename MT>
__device__ void adjust_order(BdfU &u, Bdf<MT> &b, int deltaq) {
  const int q = u.q;
  double l[QMAX + 1];
  if ((q == 2) && (deltaq != 1)) return;
  if (deltaq == 1) {
#pragma unroll
    for (int i = 0; i <= QMAX; i++) l[i] = 0.0;
    l[2] = 1.0;
    double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0;
    const double hscale = u.hscale;
    double hsum = hscale;
#pragma unroll
    for (int j = 1; j < QMAX; j++) {
      if (j < q) {
        hsum += u.tau[j + 1];
        double xi = hsum / hscale;
        prod *= xi;
        alpha0 -= 1.0 / (j + 1);
        alpha1 += 1.0 / xi;
#pragma unroll
        for (int i = QMAX; i >= 2; i--)
          if (i <= j + 2) l[i] = l[i] * xiold + l[i - 1];
        xiold = xi;
      }
    }
    double A1 = (-alpha0 - alpha1) / prod;
    // zn[L] = A1 * zn[QMAX] (saved acor); zn[j] += l[j] * zn[L], j = 2..q
    double zl = A1 * b.zn[QMAX];
    const int L = u.L;
#pragma unroll
    for (int k = 0; k <= QMAX; k++)
      if (k == L) b.zn[k] = zl;
#pragma unroll
    for (int j = 2; j <= QMAX; j++)
      if (j <= q) b.zn[j] += l[j] * zl;
  } else {
#pragma unroll
    for (int i = 0; i <= QMAX; i++) l[i] = 0.0;
    l[2] = 1.0;
    double hsum = 0.0;
    const double hscale = u.hscale;
#pragma unroll
    for (int j = 1; j <= QMAX - 2; j++) {
      if (j <= q - 2) {
        hsum += u.tau[j];
        double xi = hsum / hscale;
#pragma unroll
        for (int i = QMAX; i >= 2; i--)
          if (i <= j + 2) l[i] = l[i] * xi + l[i - 1];
      }
    }
    double zq = selz(b.zn, q);
#pragma unroll
    for (int j = 2; j < QMAX; j++)
      if (j < q) b.zn[j] -= l[j] * zq;
  }
}

// local error test; returns 0 (pass), TRY_AGAIN or a CVODE failure flag
template <typename MT>
__device__ int do_error_test(Ctx<MT> &c, BdfU &u, Bdf<MT> &b, int &nflag, double saved_t, int &nef, double &dsm) {
  dsm = u.acnrm * u.tq[2];
  if (dsm <= 1.0) return 0;
  nef++;
  nflag = PREV_ERR_FAIL;
  const int q = u.q;
  restore(u, b, saved_t);  // updates u.tn (identical in all lanes)
  wsync();
  if (c.ln.L == 0) u.netf++;
  if (nef == MXNEF) { wsync(); return -3; }  // CV_ERR_FAILURE
  double eta;
  int ret = TRY_AGAIN;
  if (nef <= MXNEF1) {
    eta = 1.0 / (pow(BIAS2 * dsm, 1.0 / u.L) + ADDON);
    eta = fmax(ETAMIN, eta);
    if (nef >= SMALL_NEF) eta = fmin(eta, ETAMXF);
    wsync();
    if (c.ln.L == 0) { u.etamax = 1.0; u.eta = eta; }
    wsync();
    rescale(u, b);
    wsync();
    return ret;
  }
  if (q > 1) {
    wsync();
    if (c.ln.L == 0) { u.etamax = 1.0; u.eta = ETAMIN; }
    wsync();
    adjust_order(u, b, -1);
    wsync();
    if (c.ln.L == 0) { u.L = q; u.q = q - 1; u.qwait = q; }
    wsync();
    rescale(u, b);
    wsync();
    return ret;
  }
  wsync();
  if (c.ln.L == 0) {
    u.etamax = 1.0;
    u.eta = ETAMIN;
    u.h *= ETAMIN;
    u.hscale = u.h;
    u.qwait = LONG_WAIT;
    u.nscon = 0;
  }
  wsync();
  double tv = f_eval(c, b.zn[0]);
  if (c.ln.L == 0) u.nfe++;
  wsync();
  b.zn[1] = u.h * tv;
  return ret;
}

template <typename MT>
__device__ void complete_step(Ctx<MT> &c, BdfU &u, Bdf<MT> &b) {
  const int q = u.q;
  const double h = u.h;
#pragma unroll
  for (int j = 0; j <= QMAX; j++)
    if (j <= q) b.zn[j] += u.l[j] * b.acor;
  const int qwait = u.qwait - 1;
  if ((qwait == 1) && (q != QMAX)) b.zn[QMAX] = b.acor;
  const double tq5 = u.tq[5];
  const int nst = u.nst + 1;
  wsync();
  if (c.ln.L == 0) {
    u.nst = nst;
    u.nscon++;
    u.hu = h;
#pragma unroll
    for (int i = QMAX; i >= 2; i--)
      if (i <= q) u.tau[i] = u.tau[i - 1];
    if ((q == 1) && (nst > 1)) u.tau[2] = u.tau[1];
    u.tau[1] = h;
    u.qwait = qwait;
    if ((qwait == 1) && (q != QMAX)) u.saved_tq5 = tq5;
  }
  wsync();
}

__device__ __forceinline__ void set_eta_vals(double &eta, double &hprime, double h, double etamax) {
  if (eta < THRESH) {
    eta = 1.0;
    hprime = h;
  } else {
    eta = fmin(eta, etamax);
    hprime = h * eta;
  }
}

template <typename MT>
__device__ void prepare_next_step(Ctx<MT> &c, BdfU &u, Bdf<MT> &b, double dsm) {
  const int q = u.q, L = u.L;
  const double h = u.h, etamax = u.etamax;
  double eta, hprime;
  int qprime = q, qwait = u.qwait;
  if (etamax == 1.0) {
    qwait = max(qwait, 2);
    hprime = h;
    eta = 1.0;
  } else {
    double etaq = 1.0 / (pow(BIAS2 * dsm, 1.0 / L) + ADDON);
    if (qwait != 0) {
      eta = etaq;
      set_eta_vals(eta, hprime, h, etamax);
    } else {
      qwait = 2;
      double etaqm1 = 0.0, etaqp1 = 0.0;
      if (q > 1) {
        double ddn = wrms(c, selz(b.zn, q), b.ewt) * u.tq[1];
        etaqm1 = 1.0 / (pow(BIAS1 * ddn, 1.0 / q) + ADDON);
      }
      const double saved_tq5 = u.saved_tq5;
      if (q != QMAX && saved_tq5 != 0.0) {
        double cquot = (u.tq[5] / saved_tq5) * pow(h / u.tau[2], (double)L);
        double dup = wrms(c, b.acor - cquot * b.zn[QMAX], b.ewt) * u.tq[3];
        etaqp1 = 1.0 / (pow(BIAS3 * dup, 1.0 / (L + 1)) + ADDON);
      }
      double etam = fmax(etaqm1, fmax(etaq, etaqp1));
      if (etam < THRESH) {
        eta = 1.0;
      } else if (etam == etaq) {
        eta = etaq;
      } else if (etam == etaqm1) {
        eta = etaqm1;
        qprime = q - 1;
      } else {
        eta = etaqp1;
        qprime = q + 1;
        b.zn[QMAX] = b.acor;
      }
      set_eta_vals(eta, hprime, h, etamax);
    }
  }
  wsync();
  if (c.ln.L == 0) {
    u.eta = eta;
    u.hprime = hprime;
    u.qprime = qprime;
    u.qwait = qwait;
    if (qprime < q && eta >= THRESH) u.nscon = 0;
  }
  wsync();
}

template <typename MT>
__device__ int bdf_step(Ctx<MT> &c, BdfU &u, Bdf<MT> &b) {
  const double saved_t = u.tn;
  double dsm = 0.0;
  int ncf = 0, nef = 0, nflag = FIRST_CALL;
  if ((u.nst > 0) && (u.hprime != u.h)) {
    const int qp = u.qprime, q = u.q;
    if (qp != q) {
      adjust_order(u, b, qp - q);
      wsync();
      if (c.ln.L == 0) { u.q = qp; u.L = qp + 1; u.qwait = qp + 1; }
      wsync();
    }
    rescale(u, b);
    wsync();
  }
  for (;;) {
    predict(u, b);
    wsync();
    set_coeffs(u);
    wsync();
    nflag = nls(c, u, b, nflag);
    if (nflag != 0) {
      restore(u, b, saved_t);
      ncf++;
      wsync();
      if (c.ln.L == 0) { u.ncfn++; u.etamax = 1.0; u.eta = ETACF; }
      wsync();
      if (ncf == MXNCF) return -4;  // CV_CONV_FAILURE
      nflag = PREV_CONV_FAIL;
      rescale(u, b);
      wsync();
      continue;
    }
    int eflag = do_error_test(c, u, b, nflag, saved_t, nef, dsm);
    if (eflag == TRY_AGAIN) continue;
    if (eflag != 0) return eflag;
    break;
  }
  complete_step(c, u, b);
  prepare_next_step(c, u, b, dsm);
  b.acor *= u.tq[2];
  const int nst = u.nst;
  wsync();
  if (c.ln.L == 0) u.etamax = (nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
  wsync();
  return 0;
}

// initial step size estimate (sets u.h); returns 0 or a failure flag
template <typename MT>
__device__ int hin(Ctx<MT> &c, BdfU &u, Bdf<MT> &b, double tout) {
  const double tn = u.tn;
  double tdiff = tout - tn;
  if (tdiff == 0.0) return -27;
  double sign = (tdiff > 0.0) ? 1.0 : -1.0;
  double tdist = fabs(tdiff);
  double tround = UROUND * fmax(fabs(tn), fabs(tout));
  if (tdist < 2.0 * tround) return -27;
  double hlb = HLB_FACTOR * tround;
  double den = HUB_FACTOR * fabs(b.zn[0]) + 1.0 / b.ewt;
  double hub_inv = wave_max(c.ln.owner ? fabs(b.zn[1]) / den : 0.0);
  double hub = HUB_FACTOR * tdist;
  if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
  double hg = sqrt(hlb * hub);
  double hres;
  if (hub < hlb) {
    hres = sign * hg;
  } else {
    double hnew = hg;
    int nfe = 0;
    for (int count1 = 1; count1 <= MAX_ITERS; count1++) {
      double hgs = hg * sign;
      double ft = f_eval(c, b.zn[0] + hgs * b.zn[1]);
      nfe++;
      double yddnrm = wrms(c, (ft - b.zn[1]) * (1.0 / hgs), b.ewt);
      hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
      if (count1 == MAX_ITERS) break;
      double hrat = hnew / hg;
      if ((hrat > 0.5) && (hrat < 2.0)) break;
      if ((count1 > 1) && (hrat > 2.0)) {
        hnew = hg;
        break;
      }
      hg = hnew;
    }
    double h0 = H_BIAS * hnew;
    if (h0 < hlb) h0 = hlb;
    if (h0 > hub) h0 = hub;
    hres = sign * h0;
    wsync();
    if (c.ln.L == 0) u.nfe += nfe;
  }
  wsync();
  if (c.ln.L == 0) u.h = hres;
  wsync();
  return 0;
}

template <typename MT>
__device__ __forceinline__ double get_dky(const BdfU &u, const Bdf<MT> &b, double t) {
  const double s = (t - u.tn) / u.h;
  const int q = u.q;
  double v = selz(b.zn, q);
#pragma unroll
  for (int j = QMAX - 1; j >= 0; j--)
    if (j < q) v = b.zn[j] + s * v;
  return v;
}

// ---------------------------------------------------------------------
// one grid point: calculate_abundances.c:132-326 semantics
// ---------------------------------------------------------------------
struct PointArgs {
  double *abundance;       // [nsys][ab_stride]
  const double *rate;      // [nsys][rate_stride]
  const double *density;   // [nsys]
  int nsys, ab_stride, rate_stride;
  double rtol, atol, t0, tend, tout0;
  int nelect0;
  long long mxstep;
  int *status;             // [nsys]
  int *stats;              // [nsys][8] or null
  unsigned int *counter;   // work queue
};

template <typename MT>
__device__ void solve_point(Ctx<MT> &c, const PointArgs &a, int sys) {
  Bdf<MT> b;
  const Lane &ln = c.ln;
  BdfU &u = c.S->u;
  // rates -> per-lane flux constants
  const double *rate = a.rate + (size_t)sys * a.rate_stride;
  const double nH = a.density[sys];
#pragma unroll
  for (int r = 0; r < NET::ROUNDS; r++) {
    int s = r * 64 + ln.L;
    int re = NET::slot_reac[s];
    uint32_t f = c.tb.fab[r];
    double k = (re >= 0) ? rate[re] : 0.0;
    c.Kp[r] = ((f >> 16) & 1) ? k * nH : k;
  }
  double *ab = a.abundance + (size_t)sys * a.ab_stride;
  const double y0 = ln.row_ok ? ab[ln.i] : 0.0;
  const double xe0 = ab[a.nelect0];
  wsync();
  c.S->X[ln.L] = (ln.L == NET::NEQ + 1) ? 1.0 : 0.0;
  if (ln.L == 0) {
    c.S->F[2 * NET::NSLOT] = 0.0;
    c.S->P[64] = 0.0;
    // CVodeInit
    u.xe_last = xe0;
    u.tn = a.t0;
    u.q = 1; u.L = 2; u.qwait = 2; u.qprime = 1;
    u.etamax = ETAMX1;
    u.hu = 0.0; u.h = 0.0; u.hscale = 0.0; u.hprime = 0.0; u.eta = 1.0;
    u.nst = u.nfe = u.nsetups = u.nje = u.nni = u.ncfn = u.netf = 0;
    u.nstlp = u.nstlj = 0;
    u.nscon = 0;
    u.saved_tq5 = 0.0;
    for (int i = 0; i <= QMAX + 1; i++) u.tau[i] = 0.0;
    for (int i = 0; i < 6; i++) u.tq[i] = 0.0;
    for (int i = 0; i <= QMAX; i++) u.l[i] = 0.0;
    u.gammap = u.gamma = u.gamrat = 0.0;
    u.rl1 = 1.0;
    u.crate = 1.0;
    u.acnrm = 0.0;
    u.jcur = 0;
  }
  wsync();
  b.zn[0] = y0;
#pragma unroll
  for (int j = 1; j <= QMAX; j++) b.zn[j] = 0.0;
  b.acor = 0.0; b.y = y0; b.ftemp = 0.0; b.wsc = 1.0; b.invp = ln.i;
#pragma unroll
  for (int j = 0; j < 16; j++) { b.J[j] = (MT)0; b.A[j] = (MT)0; }
  b.ewt = 1.0 / (c.rtol * fabs(y0) + c.atol);

  double tout = a.tout0;
  double t = a.t0;
  double yout = y0;
  int nfails = 0, status = 0, flag = 0;
  bool first = true;
  do {
    // ---- CVode(tout, CV_NORMAL)
    flag = 0;
    bool done = false;
    if (first) {
      b.zn[1] = f_eval(c, b.zn[0]);
      wsync();
      if (ln.L == 0) u.nfe++;
      wsync();
      int hf = hin(c, u, b, tout);
      if (hf != 0) {
        flag = hf;
        done = true;
        t = u.tn;
        yout = b.zn[0];
      } else {
        const double h = u.h;
        wsync();
        if (ln.L == 0) { u.hscale = h; u.hprime = h; }
        wsync();
        b.zn[1] *= h;
        first = false;
      }
    } else if ((u.tn - tout) * u.h >= 0.0) {
      t = tout;
      yout = get_dky(u, b, tout);
      done = true;
    }
    long long nstloc = 0;
    while (!done) {
      if (u.nst > 0) b.ewt = 1.0 / (c.rtol * fabs(b.zn[0]) + c.atol);
      if (nstloc >= a.mxstep) {
        flag = -1; t = u.tn; yout = b.zn[0]; break;
      }
      double nrm = wrms(c, b.zn[0], b.ewt);
      if (UROUND * nrm > 1.0) {
        flag = -2; t = u.tn; yout = b.zn[0]; break;
      }
      int kflag = bdf_step(c, u, b);
      if (kflag != 0) {
        flag = kflag; t = u.tn; yout = b.zn[0]; break;
      }
      nstloc++;
      if ((u.tn - tout) * u.h >= 0.0) {
        t = tout;
        yout = get_dky(u, b, tout);
        break;
      }
    }
    // ---- calculate_abundances.c retry / abort policy
    if (flag != 0) {
      nfails++;
      if (flag == -1 || flag == -4) {
        if (nfails >= 5) status = flag;
      } else {
        status = flag;
        nfails = 5;
      }
    }
    tout = tout * 10;
    if (tout > a.tend) tout = a.tend;
  } while (t < tout && nfails < 5);

  // write-back with clip (calculate_abundances.c:308-317)
  if (ln.owner) {
    double v = yout;
    if (v < c.atol) v = 0.0;
    ab[ln.i] = v;
  }
  wsync();
  if (ln.L == 0) {
    ab[a.nelect0] = u.xe_last;
    a.status[sys] = status;
    if (a.stats) {
      int *st = a.stats + (size_t)sys * 8;
      st[0] = u.nst; st[1] = u.nfe; st[2] = u.nje; st[3] = u.nsetups;
      st[4] = u.nni; st[5] = u.ncfn; st[6] = u.netf; st[7] = flag;
    }
  }
  wsync();
}

// persistent kernel: every wavefront pulls grid points from a queue
template <typename MT>
__global__ void __launch_bounds__(256) bdf_wave_kernel(PointArgs a) {
  __shared__ BlockLDS blk;
  for (int k = threadIdx.x; k < NET::TM * 64; k += blockDim.x) blk.rterm[k] = NET::rhs_term[k];
  __syncthreads();
  Ctx<MT> c;
  c.ln.L = lane_id();
  c.ln.i = c.ln.L & 31;
  c.ln.hh = c.ln.L >> 5;
  c.ln.row_ok = c.ln.i < NET::NEQ;
  c.ln.owner = c.ln.row_ok && c.ln.hh == 0;
  c.S = &blk.w[threadIdx.x >> 6];
  c.rtol = a.rtol;
  c.atol = a.atol;
#pragma unroll
  for (int r = 0; r < NET::ROUNDS; r++) c.tb.fab[r] = NET::slot_fab[r * 64 + c.ln.L];
  c.tb.rbase = NET::row_base[c.ln.L];
  c.tb.rnl = NET::row_nlanes[c.ln.L];
  c.tb.q = (double)NET::charge[c.ln.i];
  c.tb.rterm = blk.rterm;
  for (;;) {
    unsigned int sys = 0;
    if (c.ln.L == 0) sys = atomicAdd(a.counter, 1u);
    sys = (unsigned int)uni((int)sys);
    if (sys >= (unsigned int)a.nsys) break;
    solve_point(c, a, (int)sys);
  }
}
