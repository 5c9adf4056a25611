/*
 * Heating terms of a grid point: CPU restatement of CALC_HEATING
 * (sub_calculate_heating.F90:6-386).  TEST INFRASTRUCTURE ONLY (see
 * pdr_oracle.h).
 *
 * heat[0..11] = photoelectric (Tielens & Hollenbach 1985, grain charge by
 * Newton iteration), PAH photoelectric, Weingartner & Draine, C ionization,
 * H2 formation, H2 photodissociation, FUV pumping, cosmic rays, turbulence,
 * chemical, gas-grain, total (the total sums the PAH, C ionization, H2
 * formation and photodissociation, FUV pumping, cosmic-ray, turbulent,
 * chemical and gas-grain terms, as the reference does).
 *
 * idx[13]: 0-based abundance slots of H, H2, C, CO, He+, H3+, H3O+, HCO+, H2+
 * and e-, then 0-based reactions NRGR, NRH2, NRCI.  The chemical heating uses
 * the reference's fixed reaction numbers per network (net 0 = reduced, 1 =
 * full, 2 = MYNETWORK: the pre-set channels the reference lists after the
 * STOP that guards them, sub_calculate_heating.F90:249-260).  Literals written without a D exponent in the reference are single
 * precision there and are used as such here.
 */
#include <math.h>

#include "pdr_oracle.h"

#define HT_KB 1.38065040e-16  /* healpix_types.F90:87 */
#define HT_EV 1.60217646e-12  /* healpix_types.F90:99 */
#define HT_PI 3.141592653589793238462643383279502884197

#define SP(x) ((double)(float)(x))

static double pe_f(double x, double delta, double gamma) { return x * x * x + delta * (x * x) - gamma; }
static double pe_ff(double x, double delta) { return 3 * (x * x) + delta * (2 * x); }

void pdr_heating(const pdr_phys *ph, int net, const int *idx, double n, double T, double Td, double uv,
                 const double *ab, const double *rate, double *heat) {
  const double xH = ab[idx[0]], xH2 = ab[idx[1]], xC = ab[idx[2]], xCO = ab[idx[3]], xHEx = ab[idx[4]];
  const double xH3x = ab[idx[5]], xH3Ox = ab[idx[6]], xHCOx = ab[idx[7]], xe = ab[idx[9]];
  const double kGR = rate[idx[10]], kH2 = rate[idx[11]], kCI = rate[idx[12]];
  const double habing = 1.68 * uv;
  /* dust photoelectric heating */
  const double deltad = 1.0, deltauv = 1.8, y = 0.1, hnud = 6.0, hnuh = 13.6;
  const double xk = HT_KB * T / (hnuh * HT_EV);
  const double xd = hnud / hnuh;
  const double gamma = 2.9e-4 * y * sqrt(T) * habing / (xe * n);
  const double delta = xk - xd + gamma;
  double x = 0.5, xx = x;
  for (int it = 1; it <= 100; it++) {
    xx = x - (pe_f(x, delta, gamma) / pe_ff(x, delta));
    if (fabs(xx - x) < 1.0e-2) break;
    x = xx;
  }
  x = xx;
  const double pe = 2.7e-25 * deltauv * deltad * n * y * habing *
                    (((1.0 - x) * (1.0 - x)) / x + xk * ((x * x) - 1.0) / (x * x)) * ph->metallicity;
  /* PAH photoelectric heating and recombination cooling */
  const double phi = 1.0, alpha = 0.944;
  const double beta = 0.735 / pow(T, SP(0.068));
  const double dp = habing * sqrt(T) / (xe * n * phi);
  const double eps = 4.87e-2 / (1.0 + 4.0e-3 * pow(dp, SP(0.73))) +
                     3.65e-2 * pow(T / 1.0e4, SP(0.7)) / (1.0 + 2.0e-4 * dp);
  const double pah_h = 1.30e-24 * eps * habing * n;
  const double pah_c = 4.65e-30 * pow(T, alpha) * pow(dp, beta) * xe * n * phi * n;
  const double pah = (pah_h - pah_c) * ph->metallicity;
  /* Weingartner & Draine */
  const double c0 = 5.72, c1 = 3.45e-2, c2 = 7.08e-3, c3 = 1.98e-2, c4 = 4.95e-1, c5 = 6.92e-1, c6 = 5.20e-1;
  const double wq = habing * sqrt(T) / (xe * n);
  const double wd = ph->metallicity * 1.0e-26 * (habing * n) * (c0 + c1 * pow(T, c4)) /
                    (1.0 + c2 * pow(wq, c5) * (1.0 + c3 * pow(wq, c6)));
  /* C ionization, H2 formation, H2 photodissociation */
  const double cion = (SP(1.0) * HT_EV) * kCI * xC * n;
  const double h2f = (SP(1.5) * HT_EV) * kGR * n * xH * n;
  const double h2pd = (SP(0.4) * HT_EV) * kH2 * xH2 * n;
  /* FUV pumping of H2 */
  const double ncr = 1.0e6 / sqrt(T) /
                     (1.6 * xH * exp(-((400.0 / T) * (400.0 / T))) + 1.4 * xH2 * exp(-(18100.0 / (T + 1200.0))));
  const double fuv = (SP(2.2) * HT_EV) * 9.0 * kH2 * xH2 * n / (1.0 + ncr / n);
  /* cosmic rays, turbulence */
  const double cr = (SP(9.4) * HT_EV) * (1.3e-17 * ph->zeta) * n * xH2;
  const double lt = 5.0;
  const double vt = ph->v_turb / 1.0e5;
  const double turb = 3.5e-28 * (vt * vt * vt) * (1.0 / lt) * n;
  /* chemical heating: the reference's reactions per network (1-based numbers) */
  double chem = 0.0;
  const double ev = HT_EV;
#define R(k) rate[(k) - 1]
  if (net == 0) {
    const double xH2x = ab[idx[8]];
    chem = xH2x * n * xe * R(216) * SP(10.9) * ev + xH2x * n * xH * R(155) * SP(0.94) * ev +
           xHCOx * n * xe * n * (R(240) * (SP(7.51) * ev)) +
           xH3x * n * xe * n * (R(217) * (SP(4.76) * ev) + R(218) * (SP(9.23) * ev)) +
           xH3Ox * n * xe * n * (R(236) * (SP(1.16) * ev) + R(237) * (SP(5.63) * ev) + R(238) * (SP(6.27) * ev)) +
           xHEx * n * xH2 * n * (R(50) * (SP(6.51) * ev) + R(170) * (SP(6.51) * ev)) +
           xHEx * n * xCO * n * (R(89) * (SP(2.22) * ev) + R(90) * (SP(2.22) * ev) + R(91) * (SP(2.22) * ev));
  } else if (net == 1) {
    chem = xHCOx * n * xe * n * (R(730) * (SP(7.51) * ev)) +
           xH3x * n * xe * n * (R(706) * (SP(4.76) * ev) + R(705) * (SP(9.23) * ev)) +
           xH3Ox * n * xe * n * (R(717) * (SP(1.16) * ev) + R(716) * (SP(5.63) * ev) + R(714) * (SP(6.27) * ev)) +
           xHEx * n * xH2 * n * (R(1227) * (SP(6.51) * ev) + R(265) * (SP(6.51) * ev)) +
           xHEx * n * xCO * n * (R(1541) * (SP(2.22) * ev));
  } else if (net == 2) {
    chem = xHCOx * n * xe * n * (R(185) * (SP(7.51) * ev)) +
           xH3x * n * xe * n * (R(173) * (SP(4.76) * ev) + R(172) * (SP(9.23) * ev)) +
           xH3Ox * n * xe * n * (R(179) * (SP(1.16) * ev) + R(178) * (SP(5.63) * ev) + R(176) * (SP(6.27) * ev)) +
           xHEx * n * xH2 * n * (R(297) * (SP(6.51) * ev) + R(67) * (SP(6.51) * ev)) +
           xHEx * n * xCO * n * (R(364) * (SP(2.22) * ev));
  }
#undef R
  /* gas-grain energy exchange */
  const double acc = 0.35 * exp(-sqrt((Td + T) / 5.0e2)) + 0.1;
  const double ngr = 1.998e-12 * n * ph->metallicity * SP(100.) / ph->g2d;
  const double cgr = HT_PI * (ph->grain_radius * ph->grain_radius);
  const double gg = 4.003e-12 * n * ngr * cgr * acc * sqrt(T) * (Td - T);
  heat[0] = pe;
  heat[1] = pah;
  heat[2] = wd;
  heat[3] = cion;
  heat[4] = h2f;
  heat[5] = h2pd;
  heat[6] = fuv;
  heat[7] = cr;
  heat[8] = turb;
  heat[9] = chem;
  heat[10] = gg;
  heat[11] = pah + cion + h2f + h2pd + fuv + cr + turb + chem + gg;
}
