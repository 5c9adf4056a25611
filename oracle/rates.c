/*
 * Rate coefficients per grid point: CPU restatement of
 * CALCULATE_REACTION_RATES (calc_reac_rates.F90:16-401) and the helpers it
 * calls (shield.F90, h2_form.F90, spline.F90).  Test infrastructure (see
 * pdr_oracle.h).
 *
 * Tables (constant data, identical by necessity to the reference's):
 *   Savage & Mathis tau(lambda)/tau(V)      modules.F90:354-367 (L_GRID, X_GRID)
 *   van Dishoeck & Black log CO shielding   3DPDR.F90:1532-1537, modules.F90:346-350
 * Physical constants KB, AU, PI: healpix_types.F90:61,87,94.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pdr_oracle.h"

#define PDR_PI 3.141592653589793238462643383279502884197
#define PDR_KB 1.38065040e-16
#define PDR_AU 1.66053878e-24

/* ---------------------------------------------------------------------- */
/* natural cubic splines (spline.F90:65-121 SPLINE, :182-277 SPLINT),      */
/* written from the standard tridiagonal formulation                      */
/* ---------------------------------------------------------------------- */

/* second derivatives m[0..n-1] of the natural cubic spline through (x,y) */
static void spline_natural(const double *x, const double *y, int n, double *m) {
  double c[64], d[64];
  m[0] = 0.0;
  c[0] = 0.0; d[0] = 0.0;
  /* forward elimination of  s*m[i-1] + 2*m[i] + (1-s)*m[i+1] = rhs_i,
   * s = (x_i - x_{i-1}) / (x_{i+1} - x_{i-1}) */
  for (int i = 1; i < n - 1; i++) {
    double s = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
    double piv = s * c[i - 1] + 2.0;
    c[i] = (s - 1.0) / piv;
    double rhs = 6.0 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1])) /
                 (x[i + 1] - x[i - 1]);
    d[i] = (rhs - s * d[i - 1]) / piv;
  }
  m[n - 1] = 0.0;
  for (int i = n - 2; i >= 0; i--) m[i] = c[i] * m[i + 1] + d[i];
}

/* cubic spline value at xv; interval chosen by bisection with the same tie
 * rule as SPLINT (x equal to a knot selects the interval ending there) */
static double spline_eval(const double *x, const double *y, const double *m, int n, double xv) {
  int lo = -1, hi = n; /* 0-based; sentinels outside the table */
  while (hi - lo != 1) {
    int mid = (hi + lo) >> 1;
    if (xv > x[mid]) lo = mid; else hi = mid;
  }
  if (lo < 0) { lo = 0; hi = 1; }
  if (lo == n - 1) { lo = n - 2; hi = n - 1; }
  double h = x[hi] - x[lo];
  double a = (x[hi] - xv) / h;
  double b = (xv - x[lo]) / h;
  return a * y[lo] + b * y[hi] + ((a * a * a - a) * m[lo] + (b * b * b - b) * m[hi]) * (h * h) / 6.0;
}

static const double L_GRID[30] = {
  910.0, 950.0, 1000.0, 1050.0, 1110.0, 1180.0, 1250.0, 1390.0, 1490.0, 1600.0,
  1700.0, 1800.0, 1900.0, 2000.0, 2100.0, 2190.0, 2300.0, 2400.0, 2500.0, 2740.0,
  3440.0, 4000.0, 4400.0, 5500.0, 7000.0, 9000.0, 12500.0, 22000.0, 34000.0, 1.0e9};
static const double X_GRID[30] = {
  5.76, 5.18, 4.65, 4.16, 3.73, 3.40, 3.11, 2.74, 2.63, 2.62,
  2.54, 2.50, 2.58, 2.78, 3.01, 3.12, 2.86, 2.58, 2.35, 2.00,
  1.58, 1.42, 1.32, 1.00, 0.75, 0.48, 0.28, 0.12, 0.05, 0.00};
static const double NCO_GRID[8] = {12.0, 13.0, 14.0, 15.0, 16.0, 17.0, 18.0, 19.0};
static const double NH2_GRID[6] = {18.0, 19.0, 20.0, 21.0, 22.0, 23.0};
/* SCO[i][j]: i over NCO_GRID, j over NH2_GRID */
static const double SCO_GRID[8][6] = {
  {0.000e+00, -8.539e-02, -1.451e-01, -4.559e-01, -1.303e+00, -3.883e+00},
  {-1.408e-02, -1.015e-01, -1.612e-01, -4.666e-01, -1.312e+00, -3.888e+00},
  {-1.099e-01, -2.104e-01, -2.708e-01, -5.432e-01, -1.367e+00, -3.936e+00},
  {-4.400e-01, -5.608e-01, -6.273e-01, -8.665e-01, -1.676e+00, -4.197e+00},
  {-1.154e+00, -1.272e+00, -1.355e+00, -1.602e+00, -2.305e+00, -4.739e+00},
  {-1.888e+00, -1.973e+00, -2.057e+00, -2.303e+00, -3.034e+00, -5.165e+00},
  {-2.760e+00, -2.818e+00, -2.902e+00, -3.146e+00, -3.758e+00, -5.441e+00},
  {-4.001e+00, -4.055e+00, -4.122e+00, -4.421e+00, -5.077e+00, -6.446e+00}};

static double X_DERIV[30];
static const double X_ZERO[30] = {0};
static double SCO_DERIV[8][6]; /* row splines along NH2 (SPLIE2) */
static int tables_ready = 0;

static void init_tables(void) {
  if (tables_ready) return;
#pragma omp critical(pdr_oracle_tables)
  {
    if (!tables_ready) {
      spline_natural(L_GRID, X_GRID, 30, X_DERIV);
      for (int i = 0; i < 8; i++) spline_natural(NH2_GRID, SCO_GRID[i], 6, SCO_DERIV[i]);
      tables_ready = 1;
    }
  }
}

/* XLAMBDA shield.F90:525-549.  With spline == 0 the second-derivative
 * table is all zero, which is what the reference's shared START flag leaves
 * in place when CO photodissociation is the first caller (netgen.py,
 * xlambda_mode). */
static double xlambda(double lambda, int spline) {
  if (lambda < L_GRID[0]) lambda = L_GRID[0];
  if (lambda > L_GRID[29]) lambda = L_GRID[29];
  return spline_eval(L_GRID, X_GRID, spline ? X_DERIV : X_ZERO, 30, lambda);
}

/* SCATTER shield.F90:449-497 */
static double scatter(double av, double lambda, int spline) {
  static const double A[6] = {1.000, 2.006, -1.438, 0.7364, -0.5076, -0.0592};
  static const double K[6] = {0.7514, 0.8490, 1.013, 1.282, 2.005, 5.832};
  double tauv = av / 1.086;
  double taul = tauv * xlambda(lambda, spline);
  double s = 0.0;
  if (taul < 1.0) {
    double e = K[0] * taul;
    if (e < 100.0) s = A[0] * exp(-e);
  } else {
    for (int i = 1; i <= 5; i++) {
      double e = K[i] * taul;
      if (e < 100.0) s = s + A[i] * exp(-e);
    }
  }
  return s;
}

/* H2SHIELD1 shield.F90:231-281 (Federman, Glassgold & Kwan 1979) */
static double h2shield1(double nh2, double dopw, double radw) {
  const double fpara = 0.5, fosc = 1.0e-2;
  double taud = nh2 * fpara * (1.497358985e-2) * fosc / dopw;
  double jd, jr;
  if (taud == 0.0) jd = 1.0;
  else if (taud < 2.0) jd = exp(-(0.666666667 * taud));
  else if (taud < 10.0) jd = 0.638 * pow(taud, -1.25);
  else if (taud < 100.0) jd = 0.505 * pow(taud, -1.15);
  else jd = 0.344 * pow(taud, -1.0667);
  if (radw == 0.0) {
    jr = 0.0;
  } else {
    double r = radw / (1.772453851 * dopw);
    double t = 3.02 * pow(r * 1.0e3, -0.064);
    double u = sqrt(taud * r) / t;
    jr = r / (t * sqrt(0.785398163 + u * u));
  }
  return jd + jr;
}

/* H2PDRATE shield.F90:27-55 */
static double h2pdrate(double k0, double g0, double av, double nh2, double v_turb, int spline) {
  const double lambda = 1000.0;
  double dopw = v_turb / (lambda * 1.0e-8);
  double radw = 8.0e7;
  return k0 * g0 * h2shield1(nh2, dopw, radw) * scatter(av, lambda, spline) / 2.0;
}

/* LBAR shield.F90:570-599 */
static double lbar(double nco, double nh2) {
  double u = log10(nco + 1.0), w = log10(nh2 + 1.0);
  double l = (5675.0 - 200.6 * w) - (571.6 - 24.09 * w) * u + (18.22 - 0.7664 * w) * u * u;
  if (l < 913.6) l = 913.6;
  if (l > 1076.1) l = 1076.1;
  return l;
}

/* COSHIELD shield.F90:375-416: bicubic spline over (log N_CO, log N_H2) */
static double coshield(double nco, double nh2) {
  double lco = log10(nco + 1.0), lh2 = log10(nh2 + 1.0);
  if (lco < NCO_GRID[0]) lco = NCO_GRID[0];
  if (lh2 < NH2_GRID[0]) lh2 = NH2_GRID[0];
  if (lco > NCO_GRID[7]) lco = NCO_GRID[7];
  if (lh2 > NH2_GRID[5]) lh2 = NH2_GRID[5];
  double col[8], colm[8];
  for (int i = 0; i < 8; i++) col[i] = spline_eval(NH2_GRID, SCO_GRID[i], SCO_DERIV[i], 6, lh2);
  spline_natural(NCO_GRID, col, 8, colm);
  double v = spline_eval(NCO_GRID, col, colm, 8, lco);
  return pow(10.0, v);
}

/* COPDRATE shield.F90:83-106 */
static double copdrate(double k0, double g0, double av, double nco, double nh2, int spline) {
  double lambda = lbar(nco, nh2);
  return k0 * g0 * coshield(nco, nh2) * scatter(av, lambda, spline) / 2.0;
}

/* CIPDRATE shield.F90:132-157 (Kamp & Bertoldi 2000) */
static double cipdrate(double k0, double g0, double av, double kav, double nci, double nh2, double tgas) {
  double tauc = kav * av + 1.1e-17 * nci + (0.9 * pow(tgas, 0.27) * pow(nh2 / 1.59e21, 0.45));
  return k0 * g0 * exp(-tauc) / 2.0;
}

/* SIPDRATE shield.F90:180-200 */
static double sipdrate(double k0, double g0, double av, double kav) {
  return k0 * g0 * exp(-(kav * av)) / 2.0;
}

/* H2_FORMATION_RATE h2_form.F90:8-113 (Cazaux & Tielens 2002, 2004) */
static double h2_formation_rate(double tg, double td, double metallicity, double g2d) {
  double vth = 1.45e5 * sqrt(tg / 1.0e2);
  double stick = 1.0 / (1.0 + 0.04 * sqrt(tg + td) + 0.2 * (tg / 1.0e2) + 0.08 * (tg / 1.0e2) * (tg / 1.0e2));
  const double flux = 1.0e-10;
  const double sil_x = 8.473e-22, gra_x = 7.908e-22;
  double eff[2];
  const double mu = 0.005, nu_h2 = 3.0e12, nu_hc = 1.3e13;
  const double es[2] = {110.0, 260.0}, eh2[2] = {320.0, 520.0}, ehp[2] = {450.0, 800.0}, ehc = 3.0e4;
  for (int g = 0; g < 2; g++) {
    double f1 = mu * flux / (2 * nu_h2 * exp(-eh2[g] / td));
    double sq = 1.0 + sqrt((ehc - es[g]) / (ehp[g] - es[g]));
    double f2 = 1.0 * sq * sq / 4.0 * exp(-es[g] / td);
    double eps = 1.0 / (1.0 + nu_hc / (2 * flux) * exp(-1.5 * ehc / td) * sq * sq);
    eff[g] = 1.0 / (1.0 + f1 + f2) * eps;
  }
  return 0.5 * vth * (sil_x * eff[0] + gra_x * eff[1]) * stick * metallicity * 100.0 / g2d;
}

static double arrhenius(const pdr_net *n, int j, double T) {
  if (n->gamma[j] < -200.0 && T < n->rtmin[j]) return 0.0;
  return n->alpha[j] * pow(T / 300.0, n->beta[j]) * exp(-(n->gamma[j] / T));
}

static double photo_generic(const pdr_net *n, int j, int nrays, const double *G, const double *Av) {
  double r = 0.0;
  for (int k = 0; k < nrays; k++) r = r + n->alpha[j] * G[k] * exp(-(n->gamma[j] * Av[k])) / 2.0;
  return r;
}

static double crphot(const pdr_net *n, const pdr_phys *ph, int j, double T) {
  return n->alpha[j] * ph->zeta * pow(T / 300.0, n->beta[j]) * n->gamma[j] / (1.0 - ph->omega);
}

/* value of a duplicate-aware reaction type for reaction j */
static double typed_value(const pdr_net *n, const pdr_phys *ph, int t, int j, double T,
                          int nrays, const double *G, const double *Av) {
  switch (t) {
    case PT_THERMAL: return arrhenius(n, j, T);
    case PT_PHOTO: return photo_generic(n, j, nrays, G, Av);
    case PT_CRP: return n->alpha[j] * ph->zeta;
    default: return crphot(n, ph, j, T);
  }
}

/* duplicate-window selection, calc_reac_rates.F90:113-144 (and the copies of
 * the same loop for photo :221-239, CRP :248-263, CRPHOT :277-294) */
static void duplicate_group(const pdr_net *n, const pdr_phys *ph, int t, int i, double T,
                            int nrays, const double *G, const double *Av, double *rate) {
  int j = i;
  for (;;) {
    int next_dup = (j + 1 < n->nreac) ? n->dup[j + 1] : 0;
    if (T <= n->rtmax[j] || next_dup < n->dup[j]) {
      rate[j] = typed_value(n, ph, t, j, T, nrays, G, Av);
      return;
    }
    rate[j] = 0.0;
    j++;
  }
}

int pdr_rates(const pdr_net *n, const pdr_phys *ph, double T, double Td, int nrays,
              const double *G, const double *Av, const double *cols, double *rate) {
  init_tables();
  int neg = 0;
  for (int i = 0; i < n->nreac; i++) rate[i] = 0.0;
  for (int i = 0; i < n->nreac; i++) {
    int t = n->rtype[i];
    switch (t) {
      case PT_THERMAL: case PT_PHOTO: case PT_CRP: case PT_CRPHOT:
        if (n->dup[i] == 0) rate[i] = typed_value(n, ph, t, i, T, nrays, G, Av);
        else if (n->dup[i] == 1) duplicate_group(n, ph, t, i, T, nrays, G, Av, rate);
        break;
      case PT_H2FORM:
        rate[i] = h2_formation_rate(T, Td, ph->metallicity, ph->g2d);
        break;
      case PT_PAH:
        rate[i] = n->alpha[i] * pow(T / 100.0, n->beta[i]) * 0.4;
        break;
      case PT_PHOTO_H2:
        for (int k = 0; k < nrays; k++)
          rate[i] = rate[i] + h2pdrate(n->alpha[i], G[k], Av[k], cols[k * 5 + 0], ph->v_turb, n->xlambda_spline);
        break;
      case PT_PHOTO_HD:
        for (int k = 0; k < nrays; k++)
          rate[i] = rate[i] + h2pdrate(n->alpha[i], G[k], Av[k], cols[k * 5 + 1], ph->v_turb, n->xlambda_spline);
        break;
      case PT_PHOTO_CO:
        for (int k = 0; k < nrays; k++)
          rate[i] = rate[i] + copdrate(n->alpha[i], G[k], Av[k], cols[k * 5 + 2], cols[k * 5 + 0],
                                      n->xlambda_spline);
        break;
      case PT_PHOTO_CI:
        for (int k = 0; k < nrays; k++)
          rate[i] = rate[i] + cipdrate(n->alpha[i], G[k], Av[k], n->gamma[i], cols[k * 5 + 3],
                                       cols[k * 5 + 0], T);
        break;
      case PT_PHOTO_SI:
        for (int k = 0; k < nrays; k++)
          rate[i] = rate[i] + sipdrate(n->alpha[i], G[k], Av[k], n->gamma[i]);
        break;
      case PT_FREEZE: case PT_ELFRZE: {
        double cion;
        if (n->beta[i] == 0.0) cion = 1.0;
        else if (n->beta[i] == 1.0) cion = 1.0 + 16.71e-4 / (ph->grain_radius * T);
        else cion = 0.0;
        rate[i] = n->alpha[i] * 4.57e4 * 2.4e-22 * sqrt(T / n->gamma[i]) * cion * 0.3;
        break;
      }
      case PT_CRH:
        rate[i] = 2.06e-3 * ph->zeta * 2.4e-22 * ((n->gamma[i] <= 1210.0) ? 1.0e5 : 0.0);
        break;
      case PT_PHOTD: {
        double y = (T < 50.0) ? 3.5e-3 : (T < 85.0) ? 4.0e-3 : (T < 100.0) ? 5.5e-3 : 7.5e-3;
        for (int k = 0; k < nrays; k++)
          rate[i] = rate[i] + 1.7e8 * G[k] * exp(-(1.8 * Av[k])) * 2.4e-22 * y;
        break;
      }
      case PT_THERM:
        rate[i] = sqrt(2.0 * 1.5e15 * PDR_KB / (PDR_PI * PDR_PI * PDR_AU) * n->alpha[i] / n->gamma[i]) *
                  exp(-(n->alpha[i] / Td));
        break;
      case PT_GRAIN:
        rate[i] = n->alpha[i];
        break;
      case PT_XR:
      default:
        rate[i] = 0.0;
        break;
    }
    /* clamps, calc_reac_rates.F90:387-396 */
    if (rate[i] < 0.0) neg = 1;
    if (rate[i] > 1.0 && !n->gflag[i]) rate[i] = 1.0;
    if (rate[i] < 1.0e-99) rate[i] = 0.0;
  }
  return neg ? -1 : 0;
}

int pdr_rates_batch(const pdr_net *n, const pdr_phys *ph, int npts, const double *T,
                    const double *Td, int nrays, const double *G, const double *Av,
                    const double *cols, double *rate) {
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
  for (int p = 0; p < npts; p++) {
    if (pdr_rates(n, ph, T[p], Td[p], nrays, G + (size_t)p * nrays, Av + (size_t)p * nrays,
                  cols + (size_t)p * nrays * 5, rate + (size_t)p * n->nreac) != 0)
      bad++;
  }
  return bad;
}
