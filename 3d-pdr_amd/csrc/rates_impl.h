// Device rate evaluation, one thread per (grid point, reaction).
// Included per network inside a namespace that defines `namespace NET`.
//
// Restates CALCULATE_REACTION_RATES (calc_reac_rates.F90:16-401) with the
// photo-rate helpers of shield.F90 (H2PDRATE :27-55, COPDRATE :83-106,
// CIPDRATE :132-157, SIPDRATE :180-200, H2SHIELD1 :231-281, COSHIELD
// :375-416, SCATTER :449-497, XLAMBDA :525-549, LBAR :570-599) and
// H2_FORMATION_RATE (h2_form.F90:8-113).  Reaction-type codes come from the
// network compiler (netgen.py); duplicate temperature windows use the
// per-member form of the selection loop (gpugen.py, dup_prev_max/dup_last).

__device__ __forceinline__ double spline_eval(const double *x, const double *y, const double *m, int n,
                                              double xv) {
  int lo = -1, hi = n;
  while (hi - lo != 1) {
    int mid = (hi + lo) >> 1;
    if (xv > x[mid]) lo = mid; else hi = mid;
  }
  if (lo < 0) { lo = 0; hi = 1; }
  if (lo == n - 1) { lo = n - 2; hi = n - 1; }
  double h = x[hi] - x[lo];
  double a = (x[hi] - xv) / h, b = (xv - x[lo]) / h;
  return a * y[lo] + b * y[hi] + ((a * a * a - a) * m[lo] + (b * b * b - b) * m[hi]) * (h * h) / 6.0;
}

__device__ __forceinline__ void spline_natural8(const double *x, const double *y, double *m) {
  double c[8], d[8];
  m[0] = 0.0; c[0] = 0.0; d[0] = 0.0;
#pragma unroll
  for (int i = 1; i < 7; i++) {
    double s = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
    double piv = s * c[i - 1] + 2.0;
    c[i] = (s - 1.0) / piv;
    double rhs = 6.0 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1])) /
                 (x[i + 1] - x[i - 1]);
    d[i] = (rhs - s * d[i - 1]) / piv;
  }
  m[7] = 0.0;
#pragma unroll
  for (int i = 6; i >= 0; i--) m[i] = c[i] * m[i + 1] + d[i];
}

__device__ __forceinline__ double xlambda(double lambda) {
  if (lambda < shieldtab::L_GRID[0]) lambda = shieldtab::L_GRID[0];
  if (lambda > shieldtab::L_GRID[29]) lambda = shieldtab::L_GRID[29];
  return spline_eval(shieldtab::L_GRID, shieldtab::X_GRID,
                     NET::XLAMBDA_SPLINE ? shieldtab::X_DERIV : shieldtab::X_ZERO, 30, lambda);
}

__device__ __forceinline__ double scatter(double av, double lambda) {
  const double A[6] = {1.000, 2.006, -1.438, 0.7364, -0.5076, -0.0592};
  const double K[6] = {0.7514, 0.8490, 1.013, 1.282, 2.005, 5.832};
  double taul = (av / 1.086) * xlambda(lambda);
  double s = 0.0;
  if (taul < 1.0) {
    double e = K[0] * taul;
    if (e < 100.0) s = A[0] * exp(-e);
  } else {
#pragma unroll
    for (int i = 1; i <= 5; i++) {
      double e = K[i] * taul;
      if (e < 100.0) s = s + A[i] * exp(-e);
    }
  }
  return s;
}

__device__ __forceinline__ double h2shield1(double nh2, double dopw, double radw) {
  double taud = nh2 * 0.5 * (1.497358985e-2) * 1.0e-2 / dopw;
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

__device__ __forceinline__ double h2pdrate(double k0, double g0, double av, double nh2, double v_turb) {
  const double lambda = 1000.0;
  double dopw = v_turb / (lambda * 1.0e-8);
  return k0 * g0 * h2shield1(nh2, dopw, 8.0e7) * scatter(av, lambda) / 2.0;
}

__device__ __forceinline__ double coshield(double nco, double nh2) {
  double lco = log10(nco + 1.0), lh2 = log10(nh2 + 1.0);
  if (lco < shieldtab::NCO_GRID[0]) lco = shieldtab::NCO_GRID[0];
  if (lh2 < shieldtab::NH2_GRID[0]) lh2 = shieldtab::NH2_GRID[0];
  if (lco > shieldtab::NCO_GRID[7]) lco = shieldtab::NCO_GRID[7];
  if (lh2 > shieldtab::NH2_GRID[5]) lh2 = shieldtab::NH2_GRID[5];
  double col[8], colm[8];
#pragma unroll
  for (int i = 0; i < 8; i++)
    col[i] = spline_eval(shieldtab::NH2_GRID, shieldtab::SCO_GRID[i], shieldtab::SCO_DERIV[i], 6, lh2);
  spline_natural8(shieldtab::NCO_GRID, col, colm);
  return pow(10.0, spline_eval(shieldtab::NCO_GRID, col, colm, 8, lco));
}

__device__ __forceinline__ double copdrate(double k0, double g0, double av, double nco, double nh2) {
  double u = log10(nco + 1.0), w = log10(nh2 + 1.0);
  double lam = (5675.0 - 200.6 * w) - (571.6 - 24.09 * w) * u + (18.22 - 0.7664 * w) * u * u;
  if (lam < 913.6) lam = 913.6;
  if (lam > 1076.1) lam = 1076.1;
  return k0 * g0 * coshield(nco, nh2) * scatter(av, lam) / 2.0;
}

__device__ __forceinline__ double h2_formation_rate(double tg, double td, double metallicity, double g2d) {
  double vth = 1.45e5 * sqrt(tg / 1.0e2);
  double stick = 1.0 / (1.0 + 0.04 * sqrt(tg + td) + 0.2 * (tg / 1.0e2) + 0.08 * (tg / 1.0e2) * (tg / 1.0e2));
  const double flux = 1.0e-10, mu = 0.005, nu_h2 = 3.0e12, nu_hc = 1.3e13, ehc = 3.0e4;
  const double es[2] = {110.0, 260.0}, eh2[2] = {320.0, 520.0}, ehp[2] = {450.0, 800.0};
  double eff[2];
#pragma unroll
  for (int g = 0; g < 2; g++) {
    double f1 = mu * flux / (2 * nu_h2 * exp(-eh2[g] / td));
    double sq = 1.0 + sqrt((ehc - es[g]) / (ehp[g] - es[g]));
    double f2 = 1.0 * sq * sq / 4.0 * exp(-es[g] / td);
    double eps = 1.0 / (1.0 + nu_hc / (2 * flux) * exp(-1.5 * ehc / td) * sq * sq);
    eff[g] = 1.0 / (1.0 + f1 + f2) * eps;
  }
  return 0.5 * vth * (8.473e-22 * eff[0] + 7.908e-22 * eff[1]) * stick * metallicity * 100.0 / g2d;
}

// rate of reaction j at one point; sets *neg if the value is negative
__device__ double rate_one(int j, double T, double Td, int nrays, const double *G, const double *Av,
                           const double *cols, const pdrchem::PhysParams &ph, int *neg) {
  const int t = NET::rtype[j];
  double r = 0.0;
  const double al = NET::alpha[j], be = NET::beta[j], ga = NET::gamma[j];
  bool sel = true;
  if (NET::dup[j] != 0 && (t == 0 || t == 3 || t == 9 || t == 10))
    sel = (T > NET::dup_prev_max[j]) && (T <= NET::rtmax[j] || NET::dup_last[j]);
  switch (t) {
    case 0:  // thermal (Arrhenius), calc_reac_rates.F90:99-144
      if (sel) r = (ga < -200.0 && T < NET::rtmin[j]) ? 0.0 : al * pow(T / 300.0, be) * exp(-(ga / T));
      break;
    case 1:  // grain H2 formation
      r = h2_formation_rate(T, Td, ph.metallicity, ph.g2d);
      break;
    case 2:  // PAH
      r = al * pow(T / 100.0, be) * 0.4;
      break;
    case 3:  // generic photo
      if (sel)
        for (int k = 0; k < nrays; k++) r = r + al * G[k] * exp(-(ga * Av[k])) / 2.0;
      break;
    case 4:  // H2 photodissociation
      for (int k = 0; k < nrays; k++) r = r + h2pdrate(al, G[k], Av[k], cols ? cols[k * 5 + 0] : 0.0, ph.v_turb);
      break;
    case 5:  // HD photodissociation
      for (int k = 0; k < nrays; k++) r = r + h2pdrate(al, G[k], Av[k], cols ? cols[k * 5 + 1] : 0.0, ph.v_turb);
      break;
    case 6:  // CO photodissociation
      for (int k = 0; k < nrays; k++)
        r = r + copdrate(al, G[k], Av[k], cols ? cols[k * 5 + 2] : 0.0, cols ? cols[k * 5 + 0] : 0.0);
      break;
    case 7:  // C photoionisation
      for (int k = 0; k < nrays; k++) {
        double nc = cols ? cols[k * 5 + 3] : 0.0, nh2 = cols ? cols[k * 5 + 0] : 0.0;
        double tauc = ga * Av[k] + 1.1e-17 * nc + (0.9 * pow(T, 0.27) * pow(nh2 / 1.59e21, 0.45));
        r = r + al * G[k] * exp(-tauc) / 2.0;
      }
      break;
    case 8:  // S photoionisation
      for (int k = 0; k < nrays; k++) r = r + al * G[k] * exp(-(ga * Av[k])) / 2.0;
      break;
    case 9:  // cosmic-ray ionisation
      if (sel) r = al * ph.zeta;
      break;
    case 10:  // CR-induced photons
      if (sel) r = al * ph.zeta * pow(T / 300.0, be) * ga / (1.0 - ph.omega);
      break;
    case 11:
    case 12: {  // freeze-out
      double cion = (be == 0.0) ? 1.0 : (be == 1.0) ? 1.0 + 16.71e-4 / (ph.grain_radius * T) : 0.0;
      r = al * 4.57e4 * 2.4e-22 * sqrt(T / ga) * cion * 0.3;
      break;
    }
    case 13:  // CR-heating desorption
      r = 2.06e-3 * ph.zeta * 2.4e-22 * ((ga <= 1210.0) ? 1.0e5 : 0.0);
      break;
    case 14: {  // photodesorption
      double y = (T < 50.0) ? 3.5e-3 : (T < 85.0) ? 4.0e-3 : (T < 100.0) ? 5.5e-3 : 7.5e-3;
      for (int k = 0; k < nrays; k++) r = r + 1.7e8 * G[k] * exp(-(1.8 * Av[k])) * 2.4e-22 * y;
      break;
    }
    case 15: {  // thermal desorption; KB, PI, AU as healpix_types.F90:61,87,94
      const double KB = 1.38065040e-16, AU = 1.66053878e-24, PI = 3.141592653589793238462643383279502884197;
      r = sqrt(2.0 * 1.5e15 * KB / (PI * PI * AU) * al / ga) * exp(-(al / Td));
      break;
    }
    case 16:  // grain mantle reaction
      r = al;
      break;
    default:  // X-ray types
      r = 0.0;
      break;
  }
  if (r < 0.0) *neg = 1;
  if (r > 1.0 && !NET::gflag[j]) r = 1.0;
  if (r < 1.0e-99) r = 0.0;
  return r;
}

__global__ void __launch_bounds__(256) rates_kernel(pdrchem::RateRequest q) {
  const long long total = (long long)q.npts * NET::NREAC;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    int p = (int)(idx / NET::NREAC);
    int j = (int)(idx - (long long)p * NET::NREAC);
    int neg = 0;
    pdrchem::PhysParams ph = q.ph;
    if (q.zeta_pts) ph.zeta = q.zeta_pts[p];
    double r = rate_one(j, q.T[p], q.Td[p], q.nrays, q.G + (size_t)p * q.nrays, q.Av + (size_t)p * q.nrays,
                        q.cols ? q.cols + (size_t)p * q.nrays * 5 : nullptr, ph, &neg);
    q.rate[idx] = r;
    if (neg) atomicAdd(q.neg, 1);
  }
}
