/*
 * Level populations, line cooling and dust temperature of one grid point
 * (test infrastructure; see pdr_oracle.h): the LVG / escape-probability
 * treatment of the reference's thermal balance, restated operation by
 * operation from
 *   find_Ccoeff.F90:1-150          collisional rates at T (FIND_CCOEFF)
 *   escape_probability.F90:3-174   optical depths, escape probabilities,
 *                                  line cooling, transition matrix
 *   solvlevpop.F90:1-178           statistical equilibrium by Gauss-Jordan
 *                                  elimination with full pivoting
 *   3DPDR.F90:1848-1898            partition function, LTE populations
 *   dust_t.F90:44-97               dust temperature from the FUV field
 * Literals the reference writes without a D exponent are single precision
 * there and are used as (double)(float) here.  Compiled without contraction
 * into fused multiply-adds, like the Fortran it follows.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pdr_oracle.h"

#define KB 1.38065040e-16  /* healpix_types.F90:87-99 */
#define C_LIGHT 2.99792458e+10
#define MP 1.67262164e-24
#define HP 6.62606896e-27
#define PC 3.08568025e+18
#define PI 3.141592653589793238462643383279502884197

#define IX(i, j, n) ((size_t)(i) + (size_t)(j) * (size_t)(n)) /* column-major, 0-based */

/* FIND_CCOEFF: C[i + j*nlev] = sum over partners of the rate interpolated
 * linearly in T times the partner's density; dens = n(H), n(H+), n(e),
 * n(He), n(H2) (cm^-3) */
void pdr_find_ccoeff(const pdr_coolant *c, double T, const double *dens, double *C) {
  const int n = c->nlev, nt = c->ntemp;
  const double h_abd = dens[0], hp_abd = dens[1], e_abd = dens[2], he_abd = dens[3], h2_abd = dens[4];
  for (size_t q = 0; q < (size_t)n * n; q++) C[q] = 0.0;
  double fpara = 0.0, fortho = 0.0;
  if (h2_abd > 0.0) {
    fpara = 1.0 / (1.0 + 9.0 * exp(-170.5 / T));
    fortho = 1.0 - fpara;
  }
  for (int p = 0; p < 7; p++) {
    const double *tt = c->temps; /* TEMPERATURES(p, k) = tt[p + 7k] */
    if (tt[p] == 0.0) continue;
    int klo = 0, khi = 0; /* 1-based as the reference */
    for (int k = 1; k <= nt; k++) {
      const double tk = tt[p + 7 * (k - 1)];
      if (tk > T) {
        klo = k - 1;
        khi = k;
        break;
      } else if (tk == 0.0) {
        klo = k - 1;
        khi = k - 1;
        break;
      }
    }
    if (khi == 0) {
      klo = nt;
      khi = nt;
    } else if (khi == 1) {
      klo = 1;
      khi = 1;
    }
    double step;
    if (klo == khi)
      step = 0.0;
    else
      step = (T - tt[p + 7 * (klo - 1)]) / (tt[p + 7 * (khi - 1)] - tt[p + 7 * (klo - 1)]);
    double abd;
    switch (p) {
      case 0: abd = h2_abd; break;
      case 1: abd = h2_abd * fpara; break;
      case 2: abd = h2_abd * fortho; break;
      case 3: abd = e_abd; break;
      case 4: abd = h_abd; break;
      case 5: abd = he_abd; break;
      default: abd = hp_abd; break;
    }
    const double *col = c->col + (size_t)p * n * n * nt;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        const double lo = col[IX(i, j, n) + (size_t)(klo - 1) * n * n];
        const double hi = col[IX(i, j, n) + (size_t)(khi - 1) * n * n];
        const double tmp = lo + (hi - lo) * step;
        C[IX(i, j, n)] = C[IX(i, j, n)] + tmp * abd;
      }
  }
}

/* escape_probability: the point's own populations pop[nlev]; per ray j the
 * chain of evaluation points: epray[j] points, populations chain_pop (rows
 * of nlev, the first row of each ray being the point itself, so ray j has
 * epray[j]+1 rows starting at row pop_off[j]) and the step lengths |epoint
 * i-1 -> i| in pc (epray[j] of them starting at step_off[j]).  pseudo_ray >= 0
 * is PSEUDO_1D: only that ray carries optical depth, the others have tau =
 * 1e50, and the escape probabilities are summed, not averaged.  Outputs (all
 * column-major nlev x nlev; tau and beta nlev x nlev x nrays, pairs i > j
 * only, untouched elsewhere): transition, line, tau, beta; returns the
 * cooling rate. */
double pdr_escape_probability(const pdr_coolant *c, double Td, double Tg, double v_turb, double density,
                              double metallicity, double g2d, int nrays, int pseudo_ray, const int *epray,
                              const long long *pop_off, const double *chain_pop, const long long *step_off,
                              const double *step, const double *pop, const double *C, double *transition,
                              double *line, double *tau, double *beta) {
  const int n = c->nlev;
  double cooling = 0.0;
  double *field = calloc((size_t)n * n, sizeof(double));
  double *tau_ij = malloc(sizeof(double) * (size_t)nrays);
  double *beta_ray = malloc(sizeof(double) * (size_t)nrays);
  for (size_t q = 0; q < (size_t)n * n; q++) line[q] = 0.0;
  const double frac2 = 1.0 / sqrt(8.0 * KB * Tg / PI / MP + v_turb * v_turb);
  for (int il = 0; il < n; il++) {
    for (int jl = 0; jl < il; jl++) {
      double beta_ij = 0.0, S_ij;
      for (int j = 0; j < nrays; j++) {
        tau_ij[j] = 0.0;
        beta_ray[j] = 0.0;
      }
      const double A = c->A[IX(il, jl, n)], nu = c->F[IX(il, jl, n)];
      const double frac1 = (A * (C_LIGHT * C_LIGHT * C_LIGHT)) / (8.0 * PI * (nu * nu * nu));
      const double tmp2 = 2.0 * HP * (nu * nu * nu) / (C_LIGHT * C_LIGHT);
      double BB = tmp2 * (1.0 / (exp(HP * nu / KB / 2.7) - 1.0));
      const double ngrain = 2.0e-12 * density * metallicity * 100. / g2d;
      const double rho_grain = 2.0;
      const double emissivity = (rho_grain * ngrain) * ((double)0.01f * ((double)1.3f * nu / 3.0e11));
      const double BB_dust = tmp2 * (1.0 / (exp(HP * nu / KB / Td) - 1.0) * emissivity);
      BB = BB + BB_dust;
      const double gi = c->g[il], gj = c->g[jl];
      if (pop[il] == 0.0) {
        S_ij = 0.0;
        beta_ij = 1.0;
        goto no_line;
      }
      const double tpop = (pop[jl] * gi) / (pop[il] * gj) - 1.0;
      if (fabs(tpop) < 1.0e-50) {
        S_ij = HP * nu * pop[il] * A / 4. / PI;
        beta_ij = 1.0;
        goto with_line;
      }
      S_ij = tmp2 / tpop;
      for (int j = 0; j < nrays; j++) {
        if (pseudo_ray >= 0 && j != pseudo_ray) {
          tau_ij[j] = 1.0e50;
        } else {
          const double *pp = chain_pop + (size_t)pop_off[j] * n;
          const double *st = step + step_off[j];
          for (int i = 1; i <= epray[j]; i++) {
            const double *p0 = pp + (size_t)(i - 1) * n, *p1 = pp + (size_t)i * n;
            const double frac3 = ((p0[jl] * gi - p0[il] * gj) + (p1[jl] * gi - p1[il] * gj)) / 2. / gj;
            const double rhs2 = st[i - 1];
            const double inc = frac1 * frac2 * frac3 * rhs2 * PC;
            tau_ij[j] = tau_ij[j] + inc;
          }
        }
        if (tau_ij[j] < -5.0)
          beta_ray[j] = (1.0 - exp(5.0)) / (-5.0);
        else if (fabs(tau_ij[j]) < 1.0e-8)
          beta_ray[j] = 1.0;
        else
          beta_ray[j] = (1.0 - exp(-tau_ij[j])) / tau_ij[j];
        tau[IX(il, jl, n) + (size_t)j * n * n] = tau_ij[j];
        beta[IX(il, jl, n) + (size_t)j * n * n] = beta_ray[j];
      }
      {
        double sum = 0.0;
        for (int j = 0; j < nrays; j++) sum = sum + beta_ray[j];
        beta_ij = (pseudo_ray >= 0) ? sum : sum / (double)nrays;
      }
    with_line:
      line[IX(il, jl, n)] = A * HP * nu * pop[il] * beta_ij * (S_ij - BB) / S_ij;
      cooling = cooling + line[IX(il, jl, n)];
    no_line:
      field[IX(il, jl, n)] = (1.0 - beta_ij) * S_ij + beta_ij * BB;
      field[IX(jl, il, n)] = field[IX(il, jl, n)];
    }
  }
  for (int il = 0; il < n; il++)
    for (int jl = 0; jl < n; jl++) {
      double t = c->A[IX(il, jl, n)] + c->B[IX(il, jl, n)] * field[IX(il, jl, n)] + C[IX(il, jl, n)];
      if (fabs(t) < 1.0e-50) t = 0.0;
      transition[IX(il, jl, n)] = t;
    }
  free(field);
  free(tau_ij);
  free(beta_ray);
  return cooling;
}

/* GAUSS_JORDAN of solvlevpop.F90 (Numerical Recipes, full pivoting) on the
 * column-major n x n matrix a and right-hand side b; returns 1 on a singular
 * matrix (the reference then returns with b partly transformed) */
static int gauss_jordan(int n, double *a, double *b) {
  int *ipiv = calloc((size_t)n, sizeof(int)), *indxr = malloc(sizeof(int) * n), *indxc = malloc(sizeof(int) * n);
  int irow = 0, icol = 0, rc = 0;
#define Aa(r, c) a[IX(r, c, n)]
  for (int i = 0; i < n && !rc; i++) {
    double big = 0.0;
    for (int j = 0; j < n && !rc; j++) {
      if (ipiv[j] != 1) {
        for (int k = 0; k < n; k++) {
          if (ipiv[k] == 0) {
            if (fabs(Aa(j, k)) >= big) {
              big = fabs(Aa(j, k));
              irow = j;
              icol = k;
            }
          } else if (ipiv[k] > 1) {
            rc = 1;
            break;
          }
        }
      }
    }
    if (rc) break;
    ipiv[icol] = ipiv[icol] + 1;
    if (irow != icol) {
      for (int l = 0; l < n; l++) {
        const double d = Aa(irow, l);
        Aa(irow, l) = Aa(icol, l);
        Aa(icol, l) = d;
      }
      const double d = b[irow];
      b[irow] = b[icol];
      b[icol] = d;
    }
    indxr[i] = irow;
    indxc[i] = icol;
    if (Aa(icol, icol) == 0.0) {
      rc = 1;
      break;
    }
    const double pivinv = 1.0 / Aa(icol, icol);
    Aa(icol, icol) = 1.0;
    for (int l = 0; l < n; l++) Aa(icol, l) = Aa(icol, l) * pivinv;
    b[icol] = b[icol] * pivinv;
    for (int ll = 0; ll < n; ll++) {
      if (ll != icol) {
        const double d = Aa(ll, icol);
        Aa(ll, icol) = 0.0;
        for (int l = 0; l < n; l++) Aa(ll, l) = Aa(ll, l) - Aa(icol, l) * d;
        b[ll] = b[ll] - b[icol] * d;
      }
    }
  }
#undef Aa
  free(ipiv);
  free(indxr);
  free(indxc);
  return rc;
}

/* solvlevpop: populations (cm^-3) of the statistical equilibrium with
 * transition rates transition[i + j*n] (from level i to level j) and the
 * conservation row sum(pop) = density; negative values set to 0.  Returns
 * 1 if the elimination met a singular matrix. */
int pdr_solvlevpop(int n, const double *transition, double density, double *solution) {
  double *A = malloc(sizeof(double) * (size_t)n * n);
  for (int i = 0; i < n; i++) {
    double out1 = 0.0;
    for (int j = 0; j < n; j++) {
      out1 = out1 + transition[IX(i, j, n)];
      A[IX(i, j, n)] = transition[IX(j, i, n)];
    }
    A[IX(i, i, n)] = -out1;
  }
  for (int i = 0; i < n; i++) {
    solution[i] = 0.0;
    A[IX(n - 1, i, n)] = 1.0e-8;
  }
  solution[n - 1] = density * 1.0e-8;
  const int rc = gauss_jordan(n, A, solution);
  for (int i = 0; i < n; i++)
    if (solution[i] < 0.0) solution[i] = 0.0;
  free(A);
  return rc;
}

/* partition function and LTE populations (3DPDR.F90:1848-1898) */
void pdr_lte_populations(const pdr_coolant *c, double T, double density, double *pop) {
  double z = 0.0;
  for (int l = 0; l < c->nlev; l++) z = z + c->g[l] * exp(-c->E[l] / KB / T);
  for (int l = 0; l < c->nlev; l++) pop[l] = density * c->g[l] * exp(-c->E[l] / KB / T) / z;
}

/* CALCULATE_DUST_TEMPERATURES for one point: uv = attenuated field (Draine),
 * rad_surface[nrays] = surface field per ray */
double pdr_dust_temperature(double uv, int nrays, const double *rad_surface) {
  const double nu0 = 2.65e15, tau100 = 1.0e-3, tcmb = 2.73;
  double t = 8.9e-11 * nu0 * (1.71 * uv) + tcmb * tcmb * tcmb * tcmb * tcmb;
  for (int j = 0; j < nrays; j++) {
    const double t0 = (double)12.2f * pow(1.71 * rad_surface[j], (double)0.2f);
    if (t0 > 0)
      t = t + ((double)0.42f - log(3.45e-2 * tau100 * t0)) * (3.45e-2 * tau100 * t0) * (t0 * t0 * t0 * t0 * t0);
  }
  t = pow(t, (double)0.2f);
  if (t < 10.0) t = 10.0;
  return t;
}
