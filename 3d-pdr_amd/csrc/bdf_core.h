// CVODE 2.7.0's variable-order BDF (Nordsieck array, fixed-leading-
// coefficient orders 1-5, local error test, step and order selection, Newton
// iteration with CVODE's convergence and divergence tests, setup and
// Jacobian refresh rules, initial step estimate, CV_NORMAL output by
// interpolation) for one stiff system integrated by a team of lanes, and the
// per-point output schedule, retry policy and write-back of
// calculate_abundances.c:132-326.
//
// Included inside a per-network namespace that defines `namespace NET`.  The
// team type T (a wavefront or a workgroup) supplies the linear algebra and
// the collective operations:
//   bool   lead()            the lane that writes the shared state
//   bool   own()             this lane's row value enters the norms
//   double sum(double)       team-wide sum, result uniform
//   double vmax(double)      team-wide max, result uniform
//   void   sync()            orders shared-memory traffic of the team
//   double rhs(double y, double &xe)            ydot of this lane's row
//   int    setup(BdfU&, State&, int convfail)   M = I - gamma J and factor
//   double solve(const BdfU&, const State&, double r)   M^-1 r, own row
//   double rtol, atol
// State holds the lane's n-vector entries (zn[0..5], ewt, acor, y, ftemp,
// wsc) plus the team's matrix registers.  The higher Nordsieck entries
// zn[2..QMAX] are touched only by the step's bookkeeping (predict, restore,
// rescale, order change, completion, interpolation), not by the Newton
// loop; a State may keep them outside registers between those uses:
// hist_in() brings them into zn[], hist_out() puts them back (both may do
// nothing).
//
// Every update of the shared BdfU record is made by the leader between two
// sync() calls, after every lane has read what it needs: lanes of different
// wavefronts do not run in lockstep.

#define LDSA __attribute__((address_space(3)))

constexpr int QMAX = 5;
constexpr double UROUND = 2.220446049250313e-16;
constexpr double ETAMX1 = 10000.0, ETAMX2 = 10.0, ETAMX3 = 10.0, ETAMXF = 0.2, ETAMIN = 0.1;
constexpr double ETACF = 0.25, ADDON = 0.000001, BIAS1 = 6.0, BIAS2 = 6.0, BIAS3 = 10.0;
constexpr double THRESH = 1.5, CRDOWN = 0.3, DGMAX = 0.3, RDIV = 2.0;
// DGMAX_J: a setup after a failed Newton iteration evaluates a new Jacobian
// unless gamma moved by DGMAX_J or more since the last setup; the test is
// written !(dgamma >= DGMAX_J) so that a non-finite gamma also forces one
// (the Newton failure then counts as a convergence failure, bounded by
// MXNCF, instead of retrying the setup with the stale Jacobian forever)
constexpr double DGMAX_J = 0.2, NLSCOEF = 0.1, HLB_FACTOR = 100.0, HUB_FACTOR = 0.1;
constexpr double H_BIAS = 0.5;
constexpr int SMALL_NST = 10, MXNCF = 10, MXNEF = 7, MXNEF1 = 3, SMALL_NEF = 2, LONG_WAIT = 10;
constexpr int NLS_MAXCOR = 3, MSBP = 20, MSBJ = 50, MAX_ITERS = 4;

enum { CONV_FAIL = 4, TRY_AGAIN = 5, FIRST_CALL = 6, PREV_CONV_FAIL = 7, PREV_ERR_FAIL = 8 };
enum { CV_NO_FAILURES = 0, CV_FAIL_BAD_J = 1, CV_FAIL_OTHER = 2 };

// uniform integrator state of one system (lives in LDS)
struct BdfU {
  double tn, h, hscale, hprime, hu, eta, etamax;
  double tau[QMAX + 2], tq[6], l[QMAX + 1];
  double rl1, gamma, gammap, gamrat, crate, acnrm, saved_tq5, xe_last;
  int q, qprime, L, qwait, nscon, jcur;
  int nst, nfe, nsetups, nje, nni, ncfn, netf, nstlp, nstlj;
  // inputs of the last coefficient evaluation (set_coeffs): order, step,
  // step history, and whether tq[1], tq[3] were evaluated for them
  double ch, ctau[QMAX + 1];
  int cq, cqw;
};

// a[k] for a runtime k without a runtime index: the empty asm keeps the
// compiler from folding the select chain back into a computed address, which
// would pin the whole array in private (scratch) memory
template <int NA>
__device__ __forceinline__ double selz(const double (&a)[NA], int k) {
  double v = a[0];
#pragma unroll
  for (int j = 1; j < NA; j++) {
    double t = a[j];
    asm("" : "+v"(t));
    v = (k == j) ? t : v;
  }
  return v;
}

// x^(1/n) for x >= 0, n >= 1, to about 1e-7 relative.  CVODE takes these
// roots only to turn error-norm ratios into step-size ratios, which it then
// clamps and applies with safety factors, so FP32 accuracy is ample.  The
// binary exponent is split off exactly (x = m 2^e, e = kn + r), leaving a
// root of m 2^r in [2^(-n), 2^n) for the hardware log2 / exp2; the library
// pow(double) costs about 30x the instructions.
// e / n (C truncation) for the orders' small n: one uniform switch over
// constant divisors instead of a general integer division
__device__ __forceinline__ int div_small(int e, int n) {
  switch (n) {
    case 1: return e;
    case 2: return e / 2;
    case 3: return e / 3;
    case 4: return e / 4;
    case 5: return e / 5;
    case 6: return e / 6;
    default: return e / n;
  }
}

__device__ __forceinline__ double root_n(double x, int n) {
  if (x == 0.0) return 0.0;
  if (!(x > 0.0) || isinf(x)) return (x < 0.0 && n > 1) ? __builtin_nan("") : x;  // as pow
  int e;
  const double m = frexp(x, &e);
  const int k = div_small(e, n), r = e - k * n;
  const float f = ((float)r + __builtin_amdgcn_logf((float)m)) * __builtin_amdgcn_rcpf((float)n);
  return ldexp((double)__builtin_amdgcn_exp2f(f), k);
}

// 1/n for the orders' small n, correctly rounded (the values a division
// gives), from a table indexed by the uniform n
__constant__ double kRecip[8] = {0.0, 1.0, 1.0 / 2, 1.0 / 3, 1.0 / 4, 1.0 / 5, 1.0 / 6, 1.0 / 7};

// x^n for a small integer n >= 0
__device__ __forceinline__ double ipow(double x, int n) {
  double p = 1.0;
  for (int i = 0; i < n; i++) p *= x;
  return p;
}

template <class T>
__device__ __forceinline__ double wrms(T &c, double v, double w) {
  double p = c.own() ? v * w : 0.0;
  return sqrt(c.sum_own_sq(p * p) / NET::NEQ);
}

template <class T>
__device__ __forceinline__ double f_eval(T &c, BdfU LDSA &u, double y) {
  double xe;
  double yd = c.rhs(y, xe);
  if (c.lead()) u.xe_last = xe;
  return yd;
}

// h <- hscale*eta, zn[j] *= eta^j
template <class T, class S>
__device__ __forceinline__ void rescale(T &c, BdfU LDSA &u, S &b) {
  const double eta = u.eta;
  const int q = uni(u.q);
  const double h = u.hscale * eta;
  double factor = eta;
  b.hist_in();
#pragma unroll
  for (int j = 1; j <= QMAX; j++) {
    if (j <= q) {
      b.zn[j] *= factor;
      factor *= eta;
    }
  }
  b.hist_out();
  c.sync();
  if (c.lead()) {
    u.h = h;
    u.hscale = h;
    u.nscon = 0;
  }
  c.sync();
}

template <class T, class S>
__device__ __forceinline__ void predict(T &c, BdfU LDSA &u, S &b) {
  const double tn = u.tn + u.h;
  const int q = uni(u.q);
  b.hist_in();
#pragma unroll
  for (int k = 1; k <= QMAX; k++)
#pragma unroll
    for (int j = QMAX; j >= k; j--)
      if (j <= q && k <= q) b.zn[j - 1] += b.zn[j];
  b.hist_out();
  c.sync();
  if (c.lead()) u.tn = tn;
  c.sync();
}

template <class T, class S>
__device__ __forceinline__ void restore(T &c, BdfU LDSA &u, S &b, double saved_t) {
  const int q = uni(u.q);
  b.hist_in();
#pragma unroll
  for (int k = 1; k <= QMAX; k++)
#pragma unroll
    for (int j = QMAX; j >= k; j--)
      if (j <= q && k <= q) b.zn[j - 1] -= b.zn[j];
  b.hist_out();
  c.sync();
  if (c.lead()) u.tn = saved_t;
  c.sync();
}

// BDF coefficients l[] and error-test constants tq[] (uniform scalar code).
// They are a function of q, h and tau[1..q] alone (tq[1] and tq[3] only
// evaluated when qwait is 1, else kept), so when those equal the inputs of
// the last evaluation -- a run of steps at one step size and order, the
// common case once h settles -- the values already in u are the ones the
// evaluation would give, bit for bit, and only gammap and gamrat are due.
template <class T>
__device__ __forceinline__ void set_coeffs(T &c, BdfU LDSA &u) {
  const int q = uni(u.q);
  const double h = u.h;
  const int qw1 = uni(u.qwait) == 1;
  {
    bool same = uni(u.cq) == q && u.ch == h && (!qw1 || uni(u.cqw) != 0);
#pragma unroll
    for (int j = 1; j <= QMAX; j++)
      if (j <= q) same = same && u.tau[j] == u.ctau[j];
    if (uni((int)same)) {
      const double gamma = u.gamma;
      const int nst = uni(u.nst);
      const double gammap = (nst == 0) ? gamma : u.gammap;
      const double gamrat = (nst > 0) ? gamma / gammap : 1.0;
      c.sync();
      if (c.lead()) {
        u.gammap = gammap;
        u.gamrat = gamrat;
      }
      c.sync();
      return;
    }
  }
  double l[QMAX + 1], tau[QMAX + 2], tq[6];
#pragma unroll
  for (int i = 0; i <= QMAX + 1; i++) tau[i] = u.tau[i];
#pragma unroll
  for (int i = 0; i < 6; i++) tq[i] = u.tq[i];
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
    alpha0 -= kRecip[q];
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
  tq[2] = fabs(A1 / (alpha0 * A2));
  tq[5] = fabs(A2 * xistar_inv / (lq * xi_inv));
  if (qw1) {
    if (q > 1) {
      double C = xistar_inv / lq;
      double A3 = alpha0 + kRecip[q];
      double A4 = alpha0_hat + xi_inv;
      double Cpinv = (1.0 - A4 + A3) / A3;
      tq[1] = fabs(C * Cpinv);
    } else {
      tq[1] = 1.0;
    }
    hsum += selz(tau, q);
    xi_inv = h / hsum;
    double A5 = alpha0 - kRecip[q + 1];
    double A6 = alpha0_hat - xi_inv;
    double Cppinv = (1.0 - A6 + A5) / A2;
    tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
  }
  tq[4] = NLSCOEF / tq[2];
  const double rl1 = 1.0 / l[1];
  const double gamma = h * rl1;
  const int nst = uni(u.nst);
  const double gammap = (nst == 0) ? gamma : u.gammap;
  const double gamrat = (nst > 0) ? gamma / gammap : 1.0;
  c.sync();
  if (c.lead()) {
#pragma unroll
    for (int i = 1; i < 6; i++) u.tq[i] = tq[i];
#pragma unroll
    for (int i = 0; i <= QMAX; i++) u.l[i] = l[i];
    u.rl1 = rl1;
    u.gamma = gamma;
    u.gammap = gammap;
    u.gamrat = gamrat;
    u.cq = q;
    u.ch = h;
#pragma unroll
    for (int j = 1; j <= QMAX; j++) u.ctau[j] = tau[j];
    u.cqw = qw1;
  }
  c.sync();
}

template <class T, class S>
__device__ __forceinline__ int newton_iter(T &c, BdfU LDSA &u, S &b) {
  int m = 0;
  double del = 0.0, delp = 0.0;
  for (;;) {
    double rhs_v = u.gamma * b.ftemp - (u.rl1 * b.zn[1] + b.acor);
    double dx = c.solve(u, b, rhs_v);
    del = wrms(c, dx, b.ewt);
    b.acor += dx;
    b.y = b.zn[0] + b.acor;
    double crate = u.crate;
    if (m > 0) crate = fmax(CRDOWN * crate, del / delp);
    const double dcon = del * fmin(1.0, crate) / u.tq[4];
    const int jcur = uni(u.jcur);
    c.sync();
    if (c.lead()) {
      u.nni++;
      u.crate = crate;
    }
    if (dcon <= 1.0) {
      // the norm of the accumulated correction is needed only here, and only
      // after more than one iteration (dcon is uniform, so the whole team
      // takes this branch together)
      const double acn = (m == 0) ? del : wrms(c, b.acor, b.ewt);
      if (c.lead()) {
        u.acnrm = acn;
        u.jcur = 0;
      }
      c.sync();
      return 0;
    }
    m++;
    if ((m == NLS_MAXCOR) || ((m >= 2) && (del > RDIV * delp))) {
      c.sync();
      return jcur ? CONV_FAIL : TRY_AGAIN;
    }
    delp = del;
    c.sync();
    b.ftemp = f_eval(c, u, b.y);
    if (c.lead()) u.nfe++;
    c.sync();
  }
}

template <class T, class S>
__device__ __forceinline__ int nls(T &c, BdfU LDSA &u, S &b, int nflag) {
  int convfail = ((nflag == FIRST_CALL) || (nflag == PREV_ERR_FAIL)) ? CV_NO_FAILURES : CV_FAIL_OTHER;
  bool callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (uni(u.nst) == 0) ||
                   (uni(u.nst) >= uni(u.nstlp) + MSBP) || (fabs(u.gamrat - 1.0) > DGMAX);
  for (;;) {
    b.ftemp = f_eval(c, u, b.zn[0]);
    if (c.lead()) u.nfe++;
    c.sync();
    if (callSetup) {
      int ier = c.setup(u, b, convfail);
      const int nst = uni(u.nst);
      const double gamma = u.gamma;
      c.sync();
      if (c.lead()) {
        u.nsetups++;
        u.gamrat = 1.0;
        u.crate = 1.0;
        u.gammap = gamma;
        u.nstlp = nst;
      }
      c.sync();
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

template <class S>
__device__ __forceinline__ void adjust_order(BdfU LDSA &u, S &b, int deltaq) {
  const int q = uni(u.q);
  double l[QMAX + 1];
  if ((q == 2) && (deltaq != 1)) return;
  b.hist_in();
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
    const int L = uni(u.L);
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
  b.hist_out();
}

// local error test; returns 0 (pass), TRY_AGAIN or a CVODE failure flag
template <class T, class S>
__device__ __forceinline__ int do_error_test(T &c, BdfU LDSA &u, S &b, int &nflag, double saved_t, int &nef,
                                             double &dsm) {
  dsm = u.acnrm * u.tq[2];
  if (dsm <= 1.0) return 0;
  nef++;
  nflag = PREV_ERR_FAIL;
  const int q = uni(u.q), L = uni(u.L);
  restore(c, u, b, saved_t);
  if (c.lead()) u.netf++;
  if (nef == MXNEF) { c.sync(); return -3; }  // CV_ERR_FAILURE
  if (nef <= MXNEF1) {
    double eta = 1.0 / (root_n(BIAS2 * dsm, L) + ADDON);
    eta = fmax(ETAMIN, eta);
    if (nef >= SMALL_NEF) eta = fmin(eta, ETAMXF);
    if (c.lead()) { u.etamax = 1.0; u.eta = eta; }
    c.sync();
    rescale(c, u, b);
    return TRY_AGAIN;
  }
  if (q > 1) {
    if (c.lead()) { u.etamax = 1.0; u.eta = ETAMIN; }
    c.sync();
    adjust_order(u, b, -1);
    c.sync();
    if (c.lead()) { u.L = q; u.q = q - 1; u.qwait = q; }
    c.sync();
    rescale(c, u, b);
    return TRY_AGAIN;
  }
  const double h = u.h * ETAMIN;
  c.sync();
  if (c.lead()) {
    u.etamax = 1.0;
    u.eta = ETAMIN;
    u.h = h;
    u.hscale = h;
    u.qwait = LONG_WAIT;
    u.nscon = 0;
  }
  c.sync();
  double tv = f_eval(c, u, b.zn[0]);
  if (c.lead()) u.nfe++;
  c.sync();
  b.zn[1] = h * tv;
  return TRY_AGAIN;
}

template <class T, class S>
__device__ __forceinline__ void complete_step(T &c, BdfU LDSA &u, S &b) {
  const int q = uni(u.q);
  const double h = u.h;
  b.hist_in();
#pragma unroll
  for (int j = 0; j <= QMAX; j++)
    if (j <= q) b.zn[j] += u.l[j] * b.acor;
  const int qwait = uni(u.qwait) - 1;
  if ((qwait == 1) && (q != QMAX)) b.zn[QMAX] = b.acor;
  b.hist_out();
  const double tq5 = u.tq[5];
  const int nst = uni(u.nst) + 1;
  c.sync();
  if (c.lead()) {
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
  c.sync();
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

template <class T, class S>
__device__ __forceinline__ void prepare_next_step(T &c, BdfU LDSA &u, S &b, double dsm) {
  const int q = uni(u.q), L = uni(u.L);
  const double h = u.h, etamax = u.etamax;
  double eta, hprime;
  int qprime = q, qwait = uni(u.qwait);
  if (etamax == 1.0) {
    qwait = max(qwait, 2);
    hprime = h;
    eta = 1.0;
  } else {
    double etaq = 1.0 / (root_n(BIAS2 * dsm, L) + ADDON);
    if (qwait != 0) {
      eta = etaq;
      set_eta_vals(eta, hprime, h, etamax);
    } else {
      qwait = 2;
      b.hist_in();
      double etaqm1 = 0.0, etaqp1 = 0.0;
      if (q > 1) {
        double ddn = wrms(c, selz(b.zn, q), b.ewt) * u.tq[1];
        etaqm1 = 1.0 / (root_n(BIAS1 * ddn, q) + ADDON);
      }
      const double saved_tq5 = u.saved_tq5;
      if (q != QMAX && saved_tq5 != 0.0) {
        double cquot = (u.tq[5] / saved_tq5) * ipow(h / u.tau[2], L);
        double dup = wrms(c, b.acor - cquot * b.zn[QMAX], b.ewt) * u.tq[3];
        etaqp1 = 1.0 / (root_n(BIAS3 * dup, L + 1) + ADDON);
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
      b.hist_out();
      set_eta_vals(eta, hprime, h, etamax);
    }
  }
  c.sync();
  if (c.lead()) {
    u.eta = eta;
    u.hprime = hprime;
    u.qprime = qprime;
    u.qwait = qwait;
    if (qprime < q && eta >= THRESH) u.nscon = 0;
  }
  c.sync();
}

template <class T, class S>
__device__ __forceinline__ int bdf_step(T &c, BdfU LDSA &u, S &b) {
  const double saved_t = u.tn;
  double dsm = 0.0;
  int ncf = 0, nef = 0, nflag = FIRST_CALL;
  if ((uni(u.nst) > 0) && (u.hprime != u.h)) {
    const int qp = uni(u.qprime), q = uni(u.q);
    if (qp != q) {
      adjust_order(u, b, qp - q);
      c.sync();
      if (c.lead()) { u.q = qp; u.L = qp + 1; u.qwait = qp + 1; }
      c.sync();
    }
    rescale(c, u, b);
  }
  for (;;) {
    predict(c, u, b);
    set_coeffs(c, u);
    nflag = nls(c, u, b, nflag);
    if (nflag != 0) {
      restore(c, u, b, saved_t);
      ncf++;
      if (c.lead()) { u.ncfn++; u.etamax = 1.0; u.eta = ETACF; }
      c.sync();
      if (ncf == MXNCF) return -4;  // CV_CONV_FAILURE
      nflag = PREV_CONV_FAIL;
      rescale(c, u, b);
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
  const int nst = uni(u.nst);
  c.sync();
  if (c.lead()) u.etamax = (nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
  c.sync();
  return 0;
}

// initial step size estimate (sets u.h); returns 0 or a failure flag
template <class T, class S>
__device__ __forceinline__ int hin(T &c, BdfU LDSA &u, S &b, double tout) {
  const double tn = u.tn;
  double tdiff = tout - tn;
  if (tdiff == 0.0) return -27;
  double sign = (tdiff > 0.0) ? 1.0 : -1.0;
  double tdist = fabs(tdiff);
  double tround = UROUND * fmax(fabs(tn), fabs(tout));
  if (tdist < 2.0 * tround) return -27;
  double hlb = HLB_FACTOR * tround;
  double den = HUB_FACTOR * fabs(b.zn[0]) + 1.0 / b.ewt;
  double hub_inv = c.vmax(c.own() ? fabs(b.zn[1]) / den : 0.0);
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
      double ft = f_eval(c, u, b.zn[0] + hgs * b.zn[1]);
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
    c.sync();
    if (c.lead()) u.nfe += nfe;
  }
  c.sync();
  if (c.lead()) u.h = hres;
  c.sync();
  return 0;
}

template <class S>
__device__ __forceinline__ double get_dky(const BdfU LDSA &u, S &b, double t) {
  const double s = (t - u.tn) / u.h;
  const int q = uni(u.q);
  b.hist_in();
  double v = selz(b.zn, q);
#pragma unroll
  for (int j = QMAX - 1; j >= 0; j--)
    if (j < q) v = b.zn[j] + s * v;
  return v;
}

// ---------------------------------------------------------------------
// arguments of one batched solve (calculate_abundances_ semantics)
// ---------------------------------------------------------------------
struct PointArgs {
  double *abundance;       // [nsys][ab_stride]
  const double *rate;      // [nsys][rate_stride]
  const double *density;   // [nsys]
  int nsys, ab_stride, rate_stride;
  double rtol, atol, t0, tend, tout0;
  const double *tend_pts;  // [nsys] per-point end time, or null: tend for all
  int nelect0;
  long long mxstep;
  int xe_policy;           // 0: x_e of the last RHS evaluation; 1: charge sum of the returned abundances
  int *status;             // [nsys]
  int *stats;              // [nsys][8] or null
  unsigned int *counter;   // work queue
  // standing queue (refill): null for a closed batch of nsys points.  Else
  // the point arrays are a ring of ring_mask + 1 points in host memory that
  // the host keeps extending while the launch runs: *published counts the
  // points handed over so far (point k at ring position k & ring_mask), and
  // *counter the points taken; a worker stops taking points once *stop is
  // set, after life_ticks of its own run or after idle_ticks without work
  // (real-time clock ticks), so every launch ends by itself.
  const unsigned int *published;
  const unsigned int *stop;
  unsigned int ring_mask;
  unsigned long long idle_ticks, life_ticks;
};

constexpr unsigned int kNoPoint = 0xffffffffu;

__device__ __forceinline__ unsigned long long rt_ticks() { return __builtin_amdgcn_s_memrealtime(); }

// next point number of a standing queue, or kNoPoint when the worker is to
// stop; `born` is the worker's start (rt_ticks).  Claims only published
// points (compare-and-swap on the counter), so a worker that stops never
// holds a point; waits with growing sleeps between polls of the host words.
__device__ __forceinline__ unsigned int refill_take(const PointArgs &a, unsigned long long born) {
  const unsigned long long t0 = rt_ticks();
  int naps = 1;
  for (;;) {
    const unsigned long long now = rt_ticks();
    if (__hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u || now - born > a.life_ticks)
      return kNoPoint;
    const unsigned int p = __hip_atomic_load(a.published, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
    unsigned int c = __hip_atomic_load(a.counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while ((int)(p - c) > 0) {
      if (__hip_atomic_compare_exchange_strong(a.counter, &c, c + 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT))
        return c;
    }
    if (now - t0 > a.idle_ticks) return kNoPoint;
    for (int k = 0; k < naps; k++) __builtin_amdgcn_s_sleep(127);
    if (naps < 16) naps *= 2;
  }
}

// a point's status goes out last: in a standing queue the host takes a
// status other than the pending value it wrote as the sign that the point's
// abundances and counters are there, so every lane's stores are made visible system-wide
// first (each wavefront of the team runs this before the status store)
__device__ __forceinline__ void refill_release(const PointArgs &a) {
  if (a.published) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
}

// CVodeInit: uniform state of a fresh integration (leader writes)
__device__ __forceinline__ void bdf_init_uniform(BdfU LDSA &u, double t0, double xe0) {
  u.xe_last = xe0;
  u.tn = t0;
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
  u.cq = 0;  // no coefficients evaluated yet
  u.cqw = 0;
  u.ch = 0.0;
  for (int i = 0; i <= QMAX; i++) u.ctau[i] = 0.0;
  u.rl1 = 1.0;
  u.crate = 1.0;
  u.acnrm = 0.0;
  u.jcur = 0;
}

// The reference's output loop (calculate_abundances.c:214-305): CVode to
// tout = tout0, 10*tout0, ... (last clamped to tend) in CV_NORMAL mode, up
// to five failures with retries on -1/-4.  b holds the initial state (zn[0],
// ewt); on return yout is the lane's final value and flag the last CVode
// return; the result is the point's status contribution.
template <class T, class S>
__device__ __forceinline__ int integrate(T &c, BdfU LDSA &u, S &b, const PointArgs &a, double tend, double &yout,
                                         int &flag) {
  double tout = a.tout0;
  double t = a.t0;
  int nfails = 0, status = 0;
  bool first = true;
  flag = 0;
  do {
    // ---- CVode(tout, CV_NORMAL)
    flag = 0;
    bool done = false;
    if (first) {
      b.zn[1] = f_eval(c, u, b.zn[0]);
      if (c.lead()) u.nfe++;
      c.sync();
      int hf = hin(c, u, b, tout);
      if (hf != 0) {
        flag = hf;
        done = true;
        t = u.tn;
        yout = b.zn[0];
      } else {
        const double h = u.h;
        c.sync();
        if (c.lead()) { u.hscale = h; u.hprime = h; }
        c.sync();
        b.zn[1] *= h;
        first = false;
      }
    } else if ((u.tn - tout) * u.h >= 0.0) {
      t = tout;
      yout = get_dky(u, b, tout);
      done = true;
    }
    // the step loop carries only the step count and the exit reason; the
    // output variables are set once it ends (keeps them out of the loop's
    // register pressure)
    int nstloc = 0, why = 0;  // why: 1 = tout reached, < 0 = failure flag
    while (!done) {
      if (uni(u.nst) > 0) b.ewt = 1.0 / (c.rtol * fabs(b.zn[0]) + c.atol);
      if (nstloc >= a.mxstep) { why = -1; break; }
      double nrm = wrms(c, b.zn[0], b.ewt);
      if (UROUND * nrm > 1.0) { why = -2; break; }
      int kflag = bdf_step(c, u, b);
      if (kflag != 0) { why = kflag; break; }
      nstloc++;
      if ((u.tn - tout) * u.h >= 0.0) { why = 1; break; }
    }
    if (why == 1) {
      t = tout;
      yout = get_dky(u, b, tout);
    } else if (why < 0) {
      flag = why;
      t = u.tn;
      yout = b.zn[0];
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
    if (tout > tend) tout = tend;
  } while (t < tout && nfails < 5);
  return status;
}
