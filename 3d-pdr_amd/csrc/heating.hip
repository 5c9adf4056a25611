// Heating terms of grid points on the device: CALC_HEATING
// (sub_calculate_heating.F90:6-386), one thread per point, for the three
// networks.  Row-major inputs as the solver's: abundances [npts*nspec],
// rates [npts*nreac]; output heat [npts*12] in the order photoelectric, PAH
// photoelectric, Weingartner & Draine, C ionization, H2 formation, H2
// photodissociation, FUV pumping, cosmic rays, turbulence, chemical,
// gas-grain, total (the total leaves out the photoelectric and Weingartner &
// Draine terms, as the reference's does).  Literals the reference writes
// without a D exponent are single precision there and are rounded to float
// here too.  Memory-bound: ~20 gathered doubles in, 12 out per point.
#include <hip/hip_runtime.h>

#include "../../include/pdrchem.h"
#include "gen/net_reduced.h"
#include "gen/wg_full.h"
#include "gen/wg_mynetwork.h"
#include "pdr_internal.h"

#pragma clang fp contract(off)

namespace pdrchem {
namespace {

constexpr double kBoltz = 1.38065040e-16;  // healpix_types.F90:87
constexpr double kEv = 1.60217646e-12;     // healpix_types.F90:99
constexpr double kPi = 3.141592653589793238462643383279502884197;

// slots of the caller's index array (the Fortran module variables NH ... NRCI)
enum { S_H, S_H2, S_C, S_CO, S_HEX, S_H3X, S_H3OX, S_HCOX, S_H2X, S_E, R_GR, R_H2, R_CI, NIDX };

// one chemical-heating channel: x_a n x_b n^(1+extra_n) k_r E, E in eV
struct ChemTerm {
  int a, b, r;  // species slots, 1-based reaction number
  float e;
  int extra_n;
};

// sub_calculate_heating.F90:228-238
constexpr ChemTerm kChemReduced[] = {
    {S_H2X, S_E, 216, 10.9f, 0},   {S_H2X, S_H, 155, 0.94f, 0},  {S_HCOX, S_E, 240, 7.51f, 1},
    {S_H3X, S_E, 217, 4.76f, 1},   {S_H3X, S_E, 218, 9.23f, 1},  {S_H3OX, S_E, 236, 1.16f, 1},
    {S_H3OX, S_E, 237, 5.63f, 1},  {S_H3OX, S_E, 238, 6.27f, 1}, {S_HEX, S_H2, 50, 6.51f, 1},
    {S_HEX, S_H2, 170, 6.51f, 1},  {S_HEX, S_CO, 89, 2.22f, 1},  {S_HEX, S_CO, 90, 2.22f, 1},
    {S_HEX, S_CO, 91, 2.22f, 1}};
// sub_calculate_heating.F90:240-247
constexpr ChemTerm kChemFull[] = {
    {S_HCOX, S_E, 730, 7.51f, 1}, {S_H3X, S_E, 706, 4.76f, 1},   {S_H3X, S_E, 705, 9.23f, 1},
    {S_H3OX, S_E, 717, 1.16f, 1}, {S_H3OX, S_E, 716, 5.63f, 1},  {S_H3OX, S_E, 714, 6.27f, 1},
    {S_HEX, S_H2, 1227, 6.51f, 1}, {S_HEX, S_H2, 265, 6.51f, 1}, {S_HEX, S_CO, 1541, 2.22f, 1}};
// sub_calculate_heating.F90:254-259: the pre-set MYNETWORK channels (the
// reference guards them with a STOP that its user comments out; the same
// reactions as the full network's, under this network's numbers)
constexpr ChemTerm kChemMynetwork[] = {
    {S_HCOX, S_E, 185, 7.51f, 1}, {S_H3X, S_E, 173, 4.76f, 1},  {S_H3X, S_E, 172, 9.23f, 1},
    {S_H3OX, S_E, 179, 1.16f, 1}, {S_H3OX, S_E, 178, 5.63f, 1}, {S_H3OX, S_E, 176, 6.27f, 1},
    {S_HEX, S_H2, 297, 6.51f, 1}, {S_HEX, S_H2, 67, 6.51f, 1},  {S_HEX, S_CO, 364, 2.22f, 1}};

struct HeatArgs {
  PhysParams ph;
  int idx[NIDX];  // 0-based
  int npts;
  const double *n, *T, *Td, *uv, *ab, *rate;
  const double *zeta_pts;  // per-point cosmic-ray ionisation rate (units of ph.zeta) or null
  double *heat;
  ChemTerm chem[13];  // the network's chemical-heating channels, passed by value
};

__device__ inline double sp(float x) { return (double)x; }

template <int NSPEC, int NREAC, int NCHEM>
__global__ __launch_bounds__(256) void heating_kernel(HeatArgs a) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < a.npts; p += gridDim.x * blockDim.x) {
    const double *x = a.ab + (size_t)p * NSPEC;
    const double *k = a.rate + (size_t)p * NREAC;
    double s[10];
#pragma unroll
    for (int i = 0; i < 10; i++) s[i] = x[a.idx[i]];
    const double n = a.n[p], T = a.T[p], Td = a.Td[p];
    const double Z = a.ph.metallicity;
    const double G0 = 1.68 * a.uv[p];  // Habing units
    const double xe_n = s[S_E] * n;
    const double sqT = sqrt(T);
    // grain photoelectric heating (Tielens & Hollenbach 1985): the grain
    // charge parameter solves x^3 + delta x^2 - gamma = 0 by Newton's method
    const double xk = kBoltz * T / (13.6 * kEv);
    const double xd = 6.0 / 13.6;
    const double gam = 2.9e-4 * 0.1 * sqT * G0 / xe_n;
    const double del = xk - xd + gam;
    double xc = 0.5, xn = xc;
    for (int it = 0; it < 100; it++) {
      xn = xc - (xc * xc * xc + del * (xc * xc) - gam) / (3 * (xc * xc) + del * (2 * xc));
      if (fabs(xn - xc) < 1.0e-2) break;
      xc = xn;
    }
    xc = xn;
    const double h_pe = 2.7e-25 * 1.8 * 1.0 * n * 0.1 * G0 *
                        (((1.0 - xc) * (1.0 - xc)) / xc + xk * ((xc * xc) - 1.0) / (xc * xc)) * Z;
    // PAH photoelectric heating less recombination cooling (Bakes & Tielens 1994)
    const double dp = G0 * sqT / (xe_n * 1.0);
    const double eff = 4.87e-2 / (1.0 + 4.0e-3 * pow(dp, sp(0.73f))) +
                       3.65e-2 * pow(T / 1.0e4, sp(0.7f)) / (1.0 + 2.0e-4 * dp);
    const double h_pah = (1.30e-24 * eff * G0 * n -
                          4.65e-30 * pow(T, 0.944) * pow(dp, 0.735 / pow(T, sp(0.068f))) * s[S_E] * n * 1.0 * n) *
                         Z;
    // Weingartner & Draine (2001)
    const double wq = G0 * sqT / xe_n;
    const double h_wd = Z * 1.0e-26 * (G0 * n) * (5.72 + 3.45e-2 * pow(T, 4.95e-1)) /
                        (1.0 + 7.08e-3 * pow(wq, 6.92e-1) * (1.0 + 1.98e-2 * pow(wq, 5.20e-1)));
    const double h_ci = (sp(1.0f) * kEv) * k[a.idx[R_CI]] * s[S_C] * n;
    const double h_h2f = (sp(1.5f) * kEv) * k[a.idx[R_GR]] * n * s[S_H] * n;
    const double k_h2 = k[a.idx[R_H2]];
    const double h_h2pd = (sp(0.4f) * kEv) * k_h2 * s[S_H2] * n;
    // FUV pumping of H2, with the critical density of the vibrational levels
    const double ncrit = 1.0e6 / sqT /
                         (1.6 * s[S_H] * exp(-((400.0 / T) * (400.0 / T))) +
                          1.4 * s[S_H2] * exp(-(18100.0 / (T + 1200.0))));
    const double h_fuv = (sp(2.2f) * kEv) * 9.0 * k_h2 * s[S_H2] * n / (1.0 + ncrit / n);
    const double zeta = a.zeta_pts ? a.zeta_pts[p] : a.ph.zeta;
    const double h_cr = (sp(9.4f) * kEv) * (1.3e-17 * zeta) * n * s[S_H2];
    const double vt = a.ph.v_turb / 1.0e5;
    const double h_turb = 3.5e-28 * (vt * vt * vt) * (1.0 / 5.0) * n;
    double h_chem = 0.0;
    for (int c = 0; c < NCHEM; c++) {
      const ChemTerm t = a.chem[c];
      const double nb = t.extra_n ? n : 1.0;
      h_chem += s[t.a] * n * s[t.b] * nb * (k[t.r - 1] * ((double)t.e * kEv));
    }
    // gas-grain energy exchange (Hollenbach & McKee 1989)
    const double acc = 0.35 * exp(-sqrt((Td + T) / 5.0e2)) + 0.1;
    const double ngr = 1.998e-12 * n * Z * sp(100.0f) / a.ph.g2d;
    const double h_gg =
        4.003e-12 * n * ngr * (kPi * (a.ph.grain_radius * a.ph.grain_radius)) * acc * sqT * (Td - T);
    double *h = a.heat + (size_t)p * 12;
    h[0] = h_pe;
    h[1] = h_pah;
    h[2] = h_wd;
    h[3] = h_ci;
    h[4] = h_h2f;
    h[5] = h_h2pd;
    h[6] = h_fuv;
    h[7] = h_cr;
    h[8] = h_turb;
    h[9] = h_chem;
    h[10] = h_gg;
    h[11] = h_pah + h_ci + h_h2f + h_h2pd + h_fuv + h_cr + h_turb + h_chem + h_gg;
  }
}

template <typename Kern, int NCHEM>
int launch(Kern kern, const ChemTerm (&tab)[NCHEM], HeatArgs &a, hipStream_t s) {
  static_assert(NCHEM <= 13, "chemical-heating table larger than the argument slot");
  for (int c = 0; c < NCHEM; c++) a.chem[c] = tab[c];
  long long nb = ((long long)a.npts + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

int launch_heating(int net, const PhysParams &ph, const int *index, int npts, const double *n, const double *T,
                   const double *Td, const double *uv, const double *ab, const double *rate, const double *zeta,
                   double *heat, hipStream_t s) {
  HeatArgs a{};
  a.zeta_pts = zeta;
  a.ph = ph;
  const int nspec = net == NET_REDUCED ? pdrnet_reduced::NSPEC
                    : net == NET_FULL  ? pdrwg_full::NSPEC
                                       : pdrwg_mynetwork::NSPEC;
  const int nreac = net == NET_REDUCED ? pdrnet_reduced::NREAC
                    : net == NET_FULL  ? pdrwg_full::NREAC
                                       : pdrwg_mynetwork::NREAC;
  for (int i = 0; i < NIDX; i++) {
    a.idx[i] = index[i] - 1;
    if (a.idx[i] < 0 || a.idx[i] >= (i < R_GR ? nspec : nreac)) return -3;
  }
  a.npts = npts;
  a.n = n;
  a.T = T;
  a.Td = Td;
  a.uv = uv;
  a.ab = ab;
  a.rate = rate;
  a.heat = heat;
  switch (net) {
    case NET_REDUCED:
      return launch(heating_kernel<pdrnet_reduced::NSPEC, pdrnet_reduced::NREAC, 13>, kChemReduced, a, s);
    case NET_FULL:
      return launch(heating_kernel<pdrwg_full::NSPEC, pdrwg_full::NREAC, 9>, kChemFull, a, s);
    case NET_MYNETWORK:
      return launch(heating_kernel<pdrwg_mynetwork::NSPEC, pdrwg_mynetwork::NREAC, 9>, kChemMynetwork, a, s);
    default:
      return -2;
  }
}

}  // namespace pdrchem

using namespace pdrchem;

extern "C" int pdrchem_heating_device_zeta(int net, const pdrchem_phys *phys, const int *index, int npts,
                                           const double *d_density, const double *d_T, const double *d_Td,
                                           const double *d_uv, const double *d_abundance, const double *d_rate,
                                           const double *d_zeta, double *d_heat, void *stream) {
  if (npts == 0) return 0;
  if (net < 0 || net >= NET_COUNT || !phys || !index || npts < 0 || !d_density || !d_T || !d_Td ||
      !d_uv || !d_abundance || !d_rate || !d_heat)
    return -1;
  const PhysParams ph{phys->zeta, phys->v_turb, phys->omega, phys->grain_radius, phys->metallicity, phys->g2d};
  return launch_heating(net, ph, index, npts, d_density, d_T, d_Td, d_uv, d_abundance, d_rate, d_zeta, d_heat,
                        (hipStream_t)stream) == 0
             ? 0
             : -1;
}

extern "C" int pdrchem_heating_device(int net, const pdrchem_phys *phys, const int *index, int npts,
                                      const double *d_density, const double *d_T, const double *d_Td,
                                      const double *d_uv, const double *d_abundance, const double *d_rate,
                                      double *d_heat, void *stream) {
  return pdrchem_heating_device_zeta(net, phys, index, npts, d_density, d_T, d_Td, d_uv, d_abundance, d_rate,
                                     nullptr, d_heat, stream);
}
