// Thermal balance of a grid cloud on the device: level populations of the
// four coolants by the escape-probability (LVG) method, line cooling, dust
// temperatures and the temperature update of the reference's outer loop
// (3DPDR.F90:590-1241).  The host loop that sequences these kernels is in
// cloud.hip (pdrchem_cloud_thermal_*).
//
// Per point and coolant (escape_probability.F90:3-174, find_Ccoeff.F90,
// solvlevpop.F90): collisional rates interpolated in T; for every pair of
// levels the optical depth along each HEALPix ray, the escape probability,
// the source function, the line cooling and the mean radiation field; the
// transition matrix A + B<J> + C; the statistical equilibrium by
// Gauss-Jordan elimination with full pivoting.
//
// Optical depth.  The reference sums, for every pair (i > j) and ray, the
// trapezoid increments frac1*frac2*frac3*step*pc along the ray's evaluation
// points, frac3 being the mean over the step's two ends of
// (n_j g_i - n_i g_j)/g_j.  That sum is linear in the populations: it equals
// frac1*frac2*pc*(g_i S_j - g_j S_i)/(2 g_j) with S_l the trapezoid column
// sum over steps of (n_l(prev) + n_l(next)) * step.  level_columns_kernel forms S
// for every level once per (point, ray) -- one wavefront per (point, ray),
// one lane per level -- so a pair costs no chain walk (820 pairs of the
// 41-level CO molecule shared one walk).  The result equals the reference's
// sum to rounding.
//
// The LVG kernel runs one wavefront per (point, coolant): the pairs are
// spread over the lanes, the transition and elimination matrices live in
// LDS, and each elimination step finds the pivot with a wavefront reduction
// that reproduces the reference's choice (the last maximum of |a| in row-major
// order among the rows and columns not yet pivoted).
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/pdrchem.h"
#include "pdr_internal.h"

#pragma clang fp contract(off)

namespace pdrchem {
namespace {

constexpr double kKB = 1.38065040e-16;  // healpix_types.F90:87-99
constexpr double kC = 2.99792458e+10;
constexpr double kMP = 1.67262164e-24;
constexpr double kHP = 6.62606896e-27;
constexpr double kPC = 3.08568025e+18;
constexpr double kPI = 3.141592653589793238462643383279502884197;

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---------------------------------------------------------------------------
// dust temperature (dust_t.F90:44-97), one thread per point
// ---------------------------------------------------------------------------
__global__ void dust_kernel(int npts, int nrays, const double *__restrict__ G, const double *__restrict__ uv,
                            double *__restrict__ Td) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const double nu0 = 2.65e15, tau100 = 1.0e-3, tcmb = 2.73;
  double t = 8.9e-11 * nu0 * (1.71 * uv[p]) + tcmb * tcmb * tcmb * tcmb * tcmb;
  for (int j = 0; j < nrays; j++) {
    const double t0 = (double)12.2f * pow(1.71 * G[(long long)p * nrays + j], (double)0.2f);
    if (t0 > 0)
      t = t + ((double)0.42f - log(3.45e-2 * tau100 * t0)) * (3.45e-2 * tau100 * t0) * (t0 * t0 * t0 * t0 * t0);
  }
  t = pow(t, (double)0.2f);
  if (t < 10.0) t = 10.0;
  Td[p] = t;
}

// ---------------------------------------------------------------------------
// LTE populations (3DPDR.F90:1848-1898) of every coolant, one thread per
// point; points with skip[p] set keep theirs
// ---------------------------------------------------------------------------
__global__ void lte_kernel(LvgTables t, int npts, const double *__restrict__ T, const double *__restrict__ ab,
                           int nspec, const double *__restrict__ rho, const unsigned char *__restrict__ skip,
                           double *__restrict__ pop) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts || (skip && skip[p])) return;
  for (int c = 0; c < t.ncool; c++) {
    const CoolantDev &cd = t.c[c];
    const double dens = ab[(long long)p * nspec + t.species[c]] * rho[p];
    double z = 0.0;
    for (int l = 0; l < cd.nlev; l++) z = z + cd.g[l] * exp(-cd.E[l] / kKB / T[p]);
    for (int l = 0; l < cd.nlev; l++)
      pop[(long long)p * t.nlevtot + cd.off + l] = dens * cd.g[l] * exp(-cd.E[l] / kKB / T[p]) / z;
  }
}

// ---------------------------------------------------------------------------
// level columns: S[p][j][l] = sum over the evaluation points i = 1..n of ray j
// of (pop_l(q_{i-1}) + pop_l(q_i)) * step_i, q_0 = p; one wavefront per
// (point, ray), one lane per level (blockDim 64, levels in chunks of 64)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) level_columns_kernel(int npts, int nrays, int ray0, int nray_sel,
                                                           const long long *__restrict__ offs,
                                                           const int *__restrict__ count, const int *__restrict__ proj,
                                                           const double *__restrict__ step,
                                                           const double *__restrict__ pop, int nlevtot,
                                                           double *__restrict__ S) {
  const long long w = blockIdx.x;
  const int p = (int)(w / nray_sel), j = ray0 + (int)(w % nray_sel);
  if (p >= npts) return;
  const long long pj = (long long)p * nrays + j;
  const long long o = offs[pj];
  const int n = count[pj];
  for (int l0 = 0; l0 < nlevtot; l0 += 64) {
    const int l = l0 + (int)threadIdx.x;
    const bool on = l < nlevtot;
    double prev = on ? pop[(long long)p * nlevtot + l] : 0.0, acc = 0.0;
    for (int i = 0; i < n; i++) {
      const int q = proj[o + i];
      const double s = step[o + i];
      const double cur = on ? pop[(long long)q * nlevtot + l] : 0.0;
      acc = acc + (prev + cur) * s;
      prev = cur;
    }
    if (on) S[pj * nlevtot + l] = acc;
  }
}

// ---------------------------------------------------------------------------
// LVG level populations of one coolant, one wavefront per point
// ---------------------------------------------------------------------------
template <int kMaxLev>
__global__ void __launch_bounds__(64) lvg_kernel(LvgTables t, LvgArgs a, int ci) {
  __shared__ double M[kMaxLev * kMaxLev];   // C, then the transition matrix
  __shared__ double Fd[kMaxLev * kMaxLev];  // mean radiation field <J_ij> (symmetric), then the elimination matrix
  __shared__ double bvec[kMaxLev], dum[kMaxLev], own[kMaxLev];
  __shared__ double part[5];
  __shared__ int ipiv[kMaxLev];
  const int p = blockIdx.x;
  if (p >= a.npts || (a.skip && a.skip[p])) return;
  const int lane = threadIdx.x;
  const CoolantDev &cd = t.c[ci];
  const int n = cd.nlev, nn = n * n;
  const double T = a.T[p], Td = a.Td[p], rho = a.rho[p];
  const double *abp = a.ab + (long long)p * a.nspec;
  if (lane < 5) part[lane] = abp[t.partner[lane]] * rho;  // n(H), n(H+), n(e), n(He), n(H2)
  if (lane < n) own[lane] = a.pop[(long long)p * t.nlevtot + cd.off + lane];
  __syncthreads();
  // collisional rates (FIND_CCOEFF), column-major C[i + j n]
  {
    const double h2 = part[4];
    double fpara = 0.0, fortho = 0.0;
    if (h2 > 0.0) {
      fpara = 1.0 / (1.0 + 9.0 * exp(-170.5 / T));
      fortho = 1.0 - fpara;
    }
    for (int e = lane; e < nn; e += 64) M[e] = 0.0;
    for (int pt = 0; pt < 7; pt++) {
      const double *tt = cd.temps;
      if (tt[pt] == 0.0) continue;
      int klo = 0, khi = 0;
      for (int k = 1; k <= cd.ntemp; k++) {
        const double tk = tt[pt + 7 * (k - 1)];
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
        klo = cd.ntemp;
        khi = cd.ntemp;
      } else if (khi == 1) {
        klo = 1;
        khi = 1;
      }
      const double st = (klo == khi) ? 0.0
                                     : (T - tt[pt + 7 * (klo - 1)]) / (tt[pt + 7 * (khi - 1)] - tt[pt + 7 * (klo - 1)]);
      double abd;
      switch (pt) {
        case 0: abd = h2; break;
        case 1: abd = h2 * fpara; break;
        case 2: abd = h2 * fortho; break;
        case 3: abd = part[2]; break;
        case 4: abd = part[0]; break;
        case 5: abd = part[3]; break;
        default: abd = part[1]; break;
      }
      const double *col = cd.col + (size_t)pt * nn * cd.ntemp;
      for (int e = lane; e < nn; e += 64) {
        const double lo = col[e + (size_t)(klo - 1) * nn], hi = col[e + (size_t)(khi - 1) * nn];
        const double tmp = lo + (hi - lo) * st;
        M[e] = M[e] + tmp * abd;
      }
    }
  }
  for (int e = lane; e < nn; e += 64) Fd[e] = 0.0;
  __syncthreads();
  // pairs (il > jl): escape probabilities, line cooling, <J>
  const double frac2 = 1.0 / sqrt(8.0 * kKB * T / kPI / kMP + a.v_turb * a.v_turb);
  const double ngrain = 2.0e-12 * rho * a.metallicity * 100. / a.g2d;
  const double *Sp = a.S + (long long)p * a.nrays * t.nlevtot + cd.off;  // S[j * nlevtot + l]
  double cool = 0.0;
  const int npair = n * (n - 1) / 2;
  for (int q = lane; q < npair; q += 64) {
    // pair q -> (il, jl), il > jl, in the order (1,0), (2,0), (2,1), (3,0), ...
    int il = (int)((1.0 + sqrt(1.0 + 8.0 * (double)q)) * 0.5);
    while (il * (il - 1) / 2 > q) il--;
    while ((il + 1) * il / 2 <= q) il++;
    const int jl = q - il * (il - 1) / 2;
    const int e = il + jl * n;
    const double A = cd.A[e], nu = cd.F[e];
    const double frac1 = (A * (kC * kC * kC)) / (8.0 * kPI * (nu * nu * nu));
    const double tmp2 = 2.0 * kHP * (nu * nu * nu) / (kC * kC);
    double BB = tmp2 * (1.0 / (exp(kHP * nu / kKB / 2.7) - 1.0));
    const double emis = (2.0 * ngrain) * ((double)0.01f * ((double)1.3f * nu / 3.0e11));
    const double bbd = tmp2 * (1.0 / (exp(kHP * nu / kKB / Td) - 1.0) * emis);
    BB = BB + bbd;
    const double gi = cd.g[il], gj = cd.g[jl];
    double beta, Sij, tau_sel = 0.0;
    bool line = true;
    if (own[il] == 0.0) {
      Sij = 0.0;
      beta = 1.0;
      line = false;
    } else {
      const double tpop = (own[jl] * gi) / (own[il] * gj) - 1.0;
      if (fabs(tpop) < 1.0e-50) {
        Sij = kHP * nu * own[il] * A / 4. / kPI;
        beta = 1.0;
      } else {
        Sij = tmp2 / tpop;
        double bsum = 0.0;
        for (int j = 0; j < a.nrays; j++) {
          double tau;
          if (a.pseudo_ray >= 0 && j != a.pseudo_ray) {
            tau = 1.0e50;
          } else {
            const double Sj = Sp[(long long)j * t.nlevtot + jl], Si = Sp[(long long)j * t.nlevtot + il];
            tau = frac1 * frac2 * kPC * ((gi * Sj - gj * Si) / (2.0 * gj));
          }
          if (j == a.tau_ray) tau_sel = tau;
          double b;
          if (tau < -5.0)
            b = (1.0 - exp(5.0)) / (-5.0);
          else if (fabs(tau) < 1.0e-8)
            b = 1.0;
          else
            b = (1.0 - exp(-tau)) / tau;
          bsum = bsum + b;
        }
        beta = (a.pseudo_ray >= 0) ? bsum : bsum / (double)a.nrays;
      }
    }
    const int k = a.line_slot(ci, il, jl);
    if (line) {
      const double lv = A * kHP * nu * own[il] * beta * (Sij - BB) / Sij;
      cool = cool + lv;
      if (k >= 0) a.lines[(long long)p * LvgArgs::kLines + k] = lv;
    } else if (k >= 0) {
      a.lines[(long long)p * LvgArgs::kLines + k] = 0.0;
    }
    if (k >= 0 && a.tau_lines) a.tau_lines[(long long)p * LvgArgs::kLines + k] = tau_sel;
    const double f = (1.0 - beta) * Sij + beta * BB;
    Fd[il + jl * n] = f;
    Fd[jl + il * n] = f;
  }
  cool = wsum(cool);
  __syncthreads();
  // transition matrix (column-major) A + B <J> + C
  for (int e = lane; e < nn; e += 64) {
    double v = cd.A[e] + cd.B[e] * Fd[e] + M[e];
    if (fabs(v) < 1.0e-50) v = 0.0;
    M[e] = v;
  }
  __syncthreads();
  // solvlevpop: elimination matrix G (row-major in Fd: G[r*n + c]),
  // G(i, j) = T(j, i), G(i, i) = -sum_j T(i, j), last row 1e-8
  for (int e = lane; e < nn; e += 64) {
    const int r = e / n, cc = e % n;
    Fd[e] = M[cc + r * n];  // T(cc, r) column-major
  }
  __syncthreads();
  if (lane < n) {
    double out1 = 0.0;
    for (int j = 0; j < n; j++) out1 = out1 + M[lane + j * n];
    Fd[lane * n + lane] = -out1;
  }
  __syncthreads();
  if (lane < n) {
    Fd[(n - 1) * n + lane] = 1.0e-8;
    bvec[lane] = (lane == n - 1) ? abp[t.species[ci]] * rho * 1.0e-8 : 0.0;
    ipiv[lane] = 0;
  }
  __syncthreads();
  double *G = Fd;
  for (int step = 0; step < n; step++) {
    // pivot: the last maximum of |G(r,c)| in row-major order over rows with
    // ipiv != 1 and columns with ipiv == 0
    double big = -1.0;
    int best = -1;
    for (int e = lane; e < nn; e += 64) {
      const int r = e / n, cc = e % n;
      if (ipiv[r] == 1 || ipiv[cc] != 0) continue;
      const double v = fabs(G[e]);
      if (v >= big) {
        big = v;
        best = e;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(big, o, 64);
      const int oe = __shfl_xor(best, o, 64);
      if (ob > big || (ob == big && oe > best)) {
        big = ob;
        best = oe;
      }
    }
    const int irow = best / n, icol = best % n;
    if (lane == 0) ipiv[icol] = ipiv[icol] + 1;
    __syncthreads();
    if (irow != icol) {
      for (int l = lane; l < n; l += 64) {
        const double d0 = G[irow * n + l];
        G[irow * n + l] = G[icol * n + l];
        G[icol * n + l] = d0;
      }
      if (lane == 0) {
        const double d0 = bvec[irow];
        bvec[irow] = bvec[icol];
        bvec[icol] = d0;
      }
    }
    __syncthreads();
    if (G[icol * n + icol] == 0.0) break;  // singular: the reference returns here
    const double pivinv = 1.0 / G[icol * n + icol];
    __syncthreads();
    for (int l = lane; l < n; l += 64) G[icol * n + l] = (l == icol ? 1.0 : G[icol * n + l]) * pivinv;
    if (lane == 0) bvec[icol] = bvec[icol] * pivinv;
    if (lane < n) dum[lane] = G[lane * n + icol];
    __syncthreads();
    for (int e = lane; e < nn; e += 64) {
      const int r = e / n, cc = e % n;
      if (r == icol) continue;
      const double base = (cc == icol) ? 0.0 : G[e];
      G[e] = base - G[icol * n + cc] * dum[r];
    }
    if (lane < n && lane != icol) bvec[lane] = bvec[lane] - bvec[icol] * dum[lane];
    __syncthreads();
  }
  // negative populations from numerical noise -> 0 (solvlevpop.F90:47-53);
  // CO_FIX averaging with the previous populations (3DPDR.F90:771-846)
  if (lane < n) {
    double s = bvec[lane];
    if (s < 0.0) s = 0.0;
    const bool fix = a.co_fix && (ci < 3 || a.co_average_co);
    if (fix) {
      if (a.levpop_iteration >= 120)
        s = own[lane];
      else if (a.levpop_iteration >= 75)
        s = (double)0.5f * (s + own[lane]);
    }
    a.solution[(long long)p * t.nlevtot + cd.off + lane] = s;
  }
  if (lane == 0) a.cool[(long long)p * 4 + ci] = cool;
}

// ---------------------------------------------------------------------------
// total cooling (3DPDR.F90:850)
// ---------------------------------------------------------------------------
__global__ void total_cool_kernel(int npts, const double *__restrict__ cool, double *__restrict__ total) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  total[p] = ((cool[p * 4] + cool[p * 4 + 1]) + cool[p * 4 + 2]) + cool[p * 4 + 3];
}

// ---------------------------------------------------------------------------
// temperature update of the thermal balance (3DPDR.F90:870-1057), one
// thread per unconverged point (compact index k -> point idx[k])
// ---------------------------------------------------------------------------
__global__ void thermal_update_kernel(ThermalArgs a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.nact) return;
  const int p = a.idx[k];
  const double *h = a.heat_c + (long long)k * 12;
  for (int i = 0; i < 12; i++) a.heat[(long long)p * 12 + i] = h[i];
  const double heat = h[11], cool = a.total_cool[p];
  const double fmean = heat - cool;
  const double fratio = 2.0 * fabs(fmean) / fabs(heat + cool);
  a.fmean[p] = fmean;
  a.fratio[p] = fratio;
  if (!a.level_conv) return;
  ThermalPoint s = a.state[p];
  double T = a.T[p];
  const double Tmin = a.Tmin, Tmax = a.Tmax;
  if (a.first_time) {
    const double dummy = T;
    if (fmean == 0) {
      s.Tlow = T;
      s.Thigh = T;
    } else if (fmean > 0) {
      s.Tlow = T;
      T = 1.3 * T;
      s.prev = 'H';
    } else if (fmean < 0) {
      s.Thigh = T;
      T = 0.7 * T;
      if (T < Tmin) T = Tmin;
      s.prev = 'C';
    }
    s.Tprev = dummy;
    if (s.Tprev < Tmin) s.Tprev = Tmin;
  } else {
    double dummy = T;
    if (fratio <= a.Fcrit) s.converged = 1;
    if (!s.binarychop) {
      if (fmean > 0 && s.prev == 'H') {
        s.Tlow = T;
        T = 1.3 * T;
        s.Thigh = T;
        s.prev = 'H';
      }
      if (fmean < 0 && s.prev == 'C') {
        s.Thigh = T;
        T = 0.7 * T;
        s.Tlow = T;
        s.prev = 'C';
        if (T < Tmin) {
          T = Tmin;
          s.Tlow = Tmin;
          s.Thigh = Tmin;
        }
      }
      if (fmean > 0 && s.prev == 'C') {
        T = (s.Thigh + s.Tlow) / 2.0;
        s.binarychop = 1;
      }
      if (fmean < 0 && s.prev == 'H') {
        T = (s.Thigh + s.Tlow) / 2.0;
        s.binarychop = 1;
      }
    } else {
      if (fmean > 0) {
        s.Tlow = T;
        T = (T + s.Thigh) / 2.0;
      }
      if (fmean < 0) {
        s.Thigh = T;
        T = (T + s.Tlow) / 2.0;
      }
    }
    if (a.temp_fix && fabs(T - s.Tprev) <= a.Tdiff && fratio > a.Fcrit) s.converged = 1;
    s.Tprev = dummy;
    if (dummy < Tmin && fmean < 0) s.converged = 1;
    if (dummy > Tmax && fmean > 0) s.converged = 1;
    if (s.converged) {
      if (dummy < Tmin) {
        s.Tprev = Tmin;
        T = Tmin;
        if (s.doleveltmin) {
          s.converged = 1;
        } else {
          s.converged = 0;
          a.level_converged[p] = 0;
          dummy = Tmin;
        }
        s.doleveltmin = 1;
      }
      if (dummy > Tmax) {
        s.Tprev = Tmax;
        T = Tmax;
      }
    }
  }
  a.T[p] = T;
  a.state[p] = s;
}

// ---------------------------------------------------------------------------
// level-population convergence (3DPDR.F90:1087-1161), one thread per point:
// per coolant, every level whose new population is at least 1e-10 times the
// coolant species' abundance must have moved by at most 1e-2 (relative)
// ---------------------------------------------------------------------------
__global__ void relch_kernel(LvgTables t, int npts, const double *__restrict__ ab, int nspec,
                             const double *__restrict__ sol, const double *__restrict__ pop,
                             unsigned char *__restrict__ level_converged, unsigned char *__restrict__ cconv,
                             unsigned *__restrict__ counts) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  bool all = true;
  for (int c = 0; c < t.ncool; c++) {
    const CoolantDev &cd = t.c[c];
    const double x = ab[(long long)p * nspec + t.species[c]];
    bool okc = true;
    for (int l = 0; l < cd.nlev; l++) {
      const double s = sol[(long long)p * t.nlevtot + cd.off + l], o = pop[(long long)p * t.nlevtot + cd.off + l];
      if (s >= x * 1.0e-10) {
        double relch;
        if (s == 0.0 && o == 0.0)
          relch = 0.0;
        else
          relch = 2.0 * fabs((s - o) / (s + o));
        if (relch > 1.0e-2) okc = false;
      }
    }
    cconv[(long long)p * 4 + c] = okc ? 1 : 0;
    if (okc) atomicAdd(&counts[1 + c], 1u);
    all = all && okc;
  }
  level_converged[p] = all ? 1 : 0;
  if (all) atomicAdd(&counts[0], 1u);
}

__global__ void count_kernel(int npts, const ThermalPoint *__restrict__ st, unsigned *__restrict__ count) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < npts && st[p].converged) atomicAdd(count, 1u);
}

__global__ void copy_kernel(long long n, const double *__restrict__ src, double *__restrict__ dst) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void init_state_kernel(int npts, const double *__restrict__ T, double Tmin, double Tmax,
                                  ThermalPoint *__restrict__ st) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  ThermalPoint s{};
  const double tg = T[p];
  s.Tprev = tg;
  s.Tlow = tg / 2.0;
  s.Thigh = tg * 1.5;
  if (s.Tlow < Tmin) s.Tlow = Tmin;
  if (s.Thigh > Tmax) s.Thigh = Tmax;
  s.prev = 0;
  st[p] = s;
}

// skip flags: converged, or level_converged too when with_levels
__global__ void skip_kernel(int npts, const ThermalPoint *__restrict__ st, const unsigned char *__restrict__ lc,
                            int with_levels, unsigned char *__restrict__ skip) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  skip[p] = (st[p].converged || (with_levels && lc[p])) ? 1 : 0;
}

inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

int launch_dust_temperature(int npts, int nrays, const double *G, const double *uv, double *Td, hipStream_t s) {
  hipLaunchKernelGGL(dust_kernel, dim3(nblk(npts, 256)), dim3(256), 0, s, npts, nrays, G, uv, Td);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_lte(const LvgTables &t, int npts, const double *T, const double *ab, int nspec, const double *rho,
               const unsigned char *skip, double *pop, hipStream_t s) {
  hipLaunchKernelGGL(lte_kernel, dim3(nblk(npts, 128)), dim3(128), 0, s, t, npts, T, ab, nspec, rho, skip, pop);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_level_columns(int npts, int nrays, int ray0, int nray_sel, const long long *offs, const int *count,
                         const int *proj, const double *step, const double *pop, int nlevtot, double *S,
                         hipStream_t s) {
  const long long nw = (long long)npts * nray_sel;
  if (nw <= 0) return 0;
  hipLaunchKernelGGL(level_columns_kernel, dim3((unsigned)nw), dim3(64), 0, s, npts, nrays, ray0, nray_sel, offs,
                     count, proj, step, pop, nlevtot, S);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_lvg(const LvgTables &t, const LvgArgs &a, int coolant, hipStream_t s) {
  if (a.npts <= 0) return 0;
  const int n = t.c[coolant].nlev;
  if (n < 2 || n > LvgTables::kMaxLev) return -1;
  if (n <= 8)
    hipLaunchKernelGGL(lvg_kernel<8>, dim3((unsigned)a.npts), dim3(64), 0, s, t, a, coolant);
  else
    hipLaunchKernelGGL(lvg_kernel<LvgTables::kMaxLev>, dim3((unsigned)a.npts), dim3(64), 0, s, t, a, coolant);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_total_cool(int npts, const double *cool, double *total, hipStream_t s) {
  hipLaunchKernelGGL(total_cool_kernel, dim3(nblk(npts, 256)), dim3(256), 0, s, npts, cool, total);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_thermal_update(const ThermalArgs &a, hipStream_t s) {
  if (a.nact <= 0) return 0;
  hipLaunchKernelGGL(thermal_update_kernel, dim3(nblk(a.nact, 256)), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_relch(const LvgTables &t, int npts, const double *ab, int nspec, const double *sol, const double *pop,
                 unsigned char *level_converged, unsigned char *cconv, unsigned *counts, hipStream_t s) {
  hipLaunchKernelGGL(relch_kernel, dim3(nblk(npts, 256)), dim3(256), 0, s, t, npts, ab, nspec, sol, pop,
                     level_converged, cconv, counts);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_count_converged(int npts, const ThermalPoint *st, unsigned *count, hipStream_t s) {
  hipLaunchKernelGGL(count_kernel, dim3(nblk(npts, 256)), dim3(256), 0, s, npts, st, count);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_copy(long long n, const double *src, double *dst, hipStream_t s) {
  if (n <= 0) return 0;
  long long nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)nb), dim3(256), 0, s, n, src, dst);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_init_thermal_state(int npts, const double *T, double Tmin, double Tmax, ThermalPoint *st, hipStream_t s) {
  hipLaunchKernelGGL(init_state_kernel, dim3(nblk(npts, 256)), dim3(256), 0, s, npts, T, Tmin, Tmax, st);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_skip_flags(int npts, const ThermalPoint *st, const unsigned char *lc, int with_levels,
                      unsigned char *skip, hipStream_t s) {
  hipLaunchKernelGGL(skip_kernel, dim3(nblk(npts, 256)), dim3(256), 0, s, npts, st, lc, with_levels, skip);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace pdrchem

// ---------------------------------------------------------------------------
// C-ABI: coolant tables and the batched level-population routines
// ---------------------------------------------------------------------------
struct pdrchem_coolant_tables {
  pdrchem::LvgTables tab{};
  std::vector<void *> bufs;
  ~pdrchem_coolant_tables() {
    for (void *b : bufs) (void)hipFree(b);
  }
};

namespace pdrchem {
const LvgTables *coolant_tables_dev(const ::pdrchem_coolant_tables *t) { return t ? &t->tab : nullptr; }
}  // namespace pdrchem

namespace {
bool slots_ok(const pdrchem::LvgTables &t, int nspec) {
  for (int k = 0; k < 4; k++)
    if (t.species[k] >= nspec) return false;
  for (int k = 0; k < 5; k++)
    if (t.partner[k] >= nspec) return false;
  return true;
}
}  // namespace

extern "C" {

pdrchem_coolant_tables *pdrchem_coolant_tables_create(const pdrchem_coolant *coolants, const int *slots) {
  using namespace pdrchem;
  if (!coolants || !slots) {
    set_last_error("pdrchem_coolant_tables_create: NULL argument");
    return nullptr;
  }
  for (int k = 0; k < 9; k++)
    if (slots[k] < 0) {
      set_last_error("pdrchem_coolant_tables_create: negative abundance slot");
      return nullptr;
    }
  auto *h = new pdrchem_coolant_tables();
  LvgTables &t = h->tab;
  t.ncool = 4;
  int off = 0;
  for (int k = 0; k < 4; k++) {
    const pdrchem_coolant &q = coolants[k];
    if (q.nlev < 2 || q.nlev > LvgTables::kMaxLev || q.ntemp < 1 || !q.energies || !q.weights || !q.A || !q.B ||
        !q.freq || !q.temperatures || !q.col) {
      set_last_error("pdrchem_coolant_tables_create: bad coolant table (2..41 levels, every array given)");
      delete h;
      return nullptr;
    }
    const size_t n = q.nlev, n2 = n * n;
    const double *src[7] = {q.energies, q.weights, q.A, q.B, q.freq, q.temperatures, q.col};
    const size_t cnt[7] = {n, n, n2, n2, n2, 7 * (size_t)q.ntemp, 7 * n2 * q.ntemp};
    double *dst[7] = {};
    for (int a = 0; a < 7; a++) {
      void *d = nullptr;
      if (hipMalloc(&d, cnt[a] * sizeof(double)) != hipSuccess ||
          hipMemcpy(d, src[a], cnt[a] * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        if (d) (void)hipFree(d);
        set_last_error("pdrchem_coolant_tables_create: device allocation or copy failed");
        delete h;
        return nullptr;
      }
      h->bufs.push_back(d);
      dst[a] = static_cast<double *>(d);
    }
    t.c[k] = CoolantDev{q.nlev, q.ntemp, off, dst[0], dst[1], dst[2], dst[3], dst[4], dst[5], dst[6]};
    off += q.nlev;
  }
  t.nlevtot = off;
  for (int k = 0; k < 4; k++) t.species[k] = slots[k];
  for (int k = 0; k < 5; k++) t.partner[k] = slots[4 + k];
  return h;
}

void pdrchem_coolant_tables_destroy(pdrchem_coolant_tables *tables) { delete tables; }

int pdrchem_coolant_tables_levels(const pdrchem_coolant_tables *tables) { return tables ? tables->tab.nlevtot : -1; }

int pdrchem_level_populations_device(const pdrchem_coolant_tables *tables, int coolant, const pdrchem_phys *phys,
                                     int npts, int nspec, int nrays, int pseudo_ray, const double *d_T,
                                     const double *d_Td, const double *d_rho, const double *d_ab,
                                     const double *d_pop, const int *d_count, const long long *d_offs,
                                     const int *d_proj, const double *d_step, int co_fix, int co_average_co,
                                     int levpop_iteration, const unsigned char *d_skip, double *d_S,
                                     double *d_solution, double *d_cool, double *d_lines, int tau_ray,
                                     double *d_tau_lines, void *stream) {
  using namespace pdrchem;
  if (!tables || !phys || coolant < -1 || coolant > 3 || npts < 0 || nrays < 1 || pseudo_ray < -1 ||
      pseudo_ray >= nrays || !slots_ok(tables->tab, nspec) || (d_tau_lines && (tau_ray < 0 || tau_ray >= nrays))) {
    set_last_error("pdrchem_level_populations_device: bad argument");
    return -1;
  }
  if (npts == 0) return 0;
  if (!d_T || !d_Td || !d_rho || !d_ab || !d_pop || !d_count || !d_offs || !d_S || !d_solution || !d_cool ||
      !d_lines) {
    set_last_error("pdrchem_level_populations_device: NULL device array");
    return -1;
  }
  const LvgTables &t = tables->tab;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (launch_level_columns(npts, nrays, pseudo_ray >= 0 ? pseudo_ray : 0, pseudo_ray >= 0 ? 1 : nrays, d_offs,
                           d_count, d_proj, d_step, d_pop, t.nlevtot, d_S, s) != 0) {
    set_last_error("pdrchem_level_populations_device: level-column launch failed");
    return -1;
  }
  LvgArgs a{};
  a.npts = npts;
  a.nspec = nspec;
  a.nrays = nrays;
  a.pseudo_ray = pseudo_ray;
  a.T = d_T;
  a.Td = d_Td;
  a.rho = d_rho;
  a.ab = d_ab;
  a.pop = d_pop;
  a.S = d_S;
  a.skip = d_skip;
  a.v_turb = phys->v_turb;
  a.metallicity = phys->metallicity;
  a.g2d = phys->g2d;
  a.co_fix = co_fix;
  a.co_average_co = co_average_co;
  a.levpop_iteration = levpop_iteration;
  a.solution = d_solution;
  a.cool = d_cool;
  a.lines = d_lines;
  a.tau_lines = d_tau_lines;
  a.tau_ray = tau_ray;
  for (int k = coolant < 0 ? 0 : coolant; k <= (coolant < 0 ? 3 : coolant); k++)
    if (launch_lvg(t, a, k, s) != 0) {
      set_last_error("pdrchem_level_populations_device: launch failed");
      return -1;
    }
  return 0;
}

int pdrchem_lte_populations_device(const pdrchem_coolant_tables *tables, int npts, int nspec, const double *d_T,
                                   const double *d_ab, const double *d_rho, const unsigned char *d_skip,
                                   double *d_pop, void *stream) {
  using namespace pdrchem;
  if (!tables || npts < 0 || !slots_ok(tables->tab, nspec) || (npts > 0 && (!d_T || !d_ab || !d_rho || !d_pop))) {
    set_last_error("pdrchem_lte_populations_device: bad argument");
    return -1;
  }
  if (npts == 0) return 0;
  if (launch_lte(tables->tab, npts, d_T, d_ab, nspec, d_rho, d_skip, d_pop, static_cast<hipStream_t>(stream)) != 0) {
    set_last_error("pdrchem_lte_populations_device: launch failed");
    return -1;
  }
  return 0;
}

int pdrchem_dust_temperature_device(int npts, int nrays, const double *d_G, const double *d_uv, double *d_Td,
                                    void *stream) {
  using namespace pdrchem;
  if (npts < 0 || nrays < 1 || (npts > 0 && (!d_G || !d_uv || !d_Td))) {
    set_last_error("pdrchem_dust_temperature_device: bad argument");
    return -1;
  }
  if (npts == 0) return 0;
  if (launch_dust_temperature(npts, nrays, d_G, d_uv, d_Td, static_cast<hipStream_t>(stream)) != 0) {
    set_last_error("pdrchem_dust_temperature_device: launch failed");
    return -1;
  }
  return 0;
}

}  // extern "C"
