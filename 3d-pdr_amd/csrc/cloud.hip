// Column densities along HEALPix rays through a regular grid, and the
// device-resident chemistry iteration over the grid (pdrchem_cloud_*).
//
// Ray tracer: the evaluation-point construction of eval_points.F90:129-225
// on a regular nx*ny*nz grid.  For an origin p the reference visits every
// other point in order of distance, finds the HEALPix ray whose pixel holds
// the direction (vec2ang + ang2pix_nest_id, healpix.F90:267-398), and accepts
// the point as the ray's next evaluation point when the angle between
// (point - ep) and the ray, seen from the ray's current evaluation point ep,
// is at most theta_crit; ep then moves to the point's projection on the ray.
// On a grid the offsets o = point - p, their pixel and their distance do not
// depend on p, so the host builds once, per ray, the list of offsets of that
// ray's pixel sorted by (|o|^2, linear offset) -- which is the order by
// (distance, point index) for every origin -- and one thread per (origin,
// ray) walks its ray's list, skipping offsets outside the grid.  Threads of
// a wavefront share the ray and read the same list entry (scalar loads).  A
// first pass counts evaluation points per (origin, ray), an exclusive scan
// gives CSR offsets, a second pass writes the projected point and the step
// length |epoint_i - epoint_{i-1}| of each evaluation point.  There is no
// maxpoints cap (the reference STOPs beyond 600, eval_points.F90:210).
//
// Columns: Av per ray (CALC_UVFIELD, 3DPDR.F90:1650-1664) and the columns of
// H2, HD, CO, C, S (calc_columndens, 3DPDR.F90:1818-1838: trapezoid of
// n_H x_k times the step, times pc) by one thread per (point, ray).
//
// Both follow the reference's expressions operation by operation, without
// fused multiply-adds, so a grid with a power-of-two spacing (exact
// coordinates and differences) gives the same evaluation points as a
// point-cloud tracer in double precision.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <thread>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pdrchem.h"
#include "pdr_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr double kPc = 3.08568025e18;  // healpix_types.F90:98
constexpr double kHalfPi = 1.570796326794896619231321691639751442099;
constexpr double kTwoPi = 6.283185307179586476925286766559005768394;
constexpr double kTwoThird = 0.6666666666666666666666666666666666666666;
constexpr int kNsMax = 8192;

thread_local std::string t_err;

bool ok(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  t_err = std::string(what) + ": " + hipGetErrorString(e);
  return false;
}

// ---------------------------------------------------------------------------
// HEALPix, nested scheme (Gorski et al. 2005, ApJ 622:759), host side
// ---------------------------------------------------------------------------
int spread_bits(int v) {
  int r = 0;
  for (int p = 0; v; p++, v >>= 1) r |= (v & 1) << (2 * p);
  return r;
}
int gather_bits(int k, int odd) {
  int r = 0;
  k >>= odd;
  for (int p = 0; k; p++, k >>= 2) r |= (k & 1) << p;
  return r;
}

// centre of pixel ipix (healpix.F90:47-223, small components set to 0 as :172-173)
void pix_centre(int nside, int ipix, double v[3]) {
  static const int jrll[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
  static const int jpll[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};
  const double fn = (double)nside;
  const double fact1 = 1.0 / (3.0 * fn * fn), fact2 = 2.0 / (3.0 * fn);
  const int nl4 = 4 * nside, npface = nside * nside;
  const int face = ipix / npface, ipf = ipix % npface;
  const int ix = gather_bits(ipf, 0), iy = gather_bits(ipf, 1);
  const int jr = jrll[face] * nside - (ix + iy) - 1;
  int nr = nside, kshift = ((jr - nside) % 2 + 2) % 2;
  double z = (2 * nside - jr) * fact2;
  if (jr < nside) {
    nr = jr;
    z = 1.0 - nr * nr * fact1;
    kshift = 0;
  } else if (jr > 3 * nside) {
    nr = nl4 - jr;
    z = -1.0 + nr * nr * fact1;
    kshift = 0;
  }
  int jp = (jpll[face] * nr + (ix - iy) + 1 + kshift) / 2;
  if (jp > nl4) jp -= nl4;
  if (jp < 1) jp += nl4;
  const double phi = (jp - (kshift + 1) * 0.5) * (kHalfPi / nr);
  const double sth = std::sqrt((1.0 - z) * (1.0 + z));
  v[0] = sth * std::cos(phi);
  v[1] = sth * std::sin(phi);
  v[2] = z;
  if (std::fabs(v[1]) < 1e-10) v[1] = 0.0;
  if (std::fabs(v[0]) < 1e-10) v[0] = 0.0;
}

// pixel holding direction r: angles first, then the pixel at resolution
// kNsMax degraded to nside (vec2ang, ang2pix_nest_id)
int pix_of(int nside, const double r[3]) {
  const double dnorm = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  double theta = 0.0, phi = 0.0;
  if (dnorm != 0.0) {
    theta = std::acos(r[2] / dnorm);
    if (r[0] != 0.0 || r[1] != 0.0) phi = std::atan2(r[1], r[0]);
    if (phi < 0.0) phi = phi + kTwoPi;
  }
  const double z = std::cos(theta), za = std::fabs(z);
  const double tt = (phi - std::floor(phi / kTwoPi) * kTwoPi) / kHalfPi;
  int face, ix, iy;
  if (za <= kTwoThird) {
    const int jp = (int)(kNsMax * (0.5 + tt - z * 0.75));
    const int jm = (int)(kNsMax * (0.5 + tt + z * 0.75));
    const int ifp = jp / kNsMax, ifm = jm / kNsMax;
    face = ifp == ifm ? (ifp % 4) + 4 : (ifp < ifm ? ifp % 4 : (ifm % 4) + 8);
    ix = jm % kNsMax;
    iy = kNsMax - (jp % kNsMax) - 1;
  } else {
    int ntt = (int)tt;
    if (ntt >= 4) ntt = 3;
    const double tp = tt - ntt, tmp = std::sqrt(3.0 * (1.0 - za));
    int jp = (int)(kNsMax * tp * tmp), jm = (int)(kNsMax * (1.0 - tp) * tmp);
    jp = std::min(kNsMax - 1, jp);
    jm = std::min(kNsMax - 1, jm);
    if (z >= 0) {
      face = ntt;
      ix = kNsMax - jm - 1;
      iy = kNsMax - jp - 1;
    } else {
      face = ntt + 8;
      ix = jp;
      iy = jm;
    }
  }
  int ipf = (spread_bits(ix / 128) + 2 * spread_bits(iy / 128)) * (128 * 128) +
            (spread_bits(ix % 128) + 2 * spread_bits(iy % 128));
  const int d = kNsMax / nside;
  return ipf / (d * d) + face * nside * nside;
}

// ---------------------------------------------------------------------------
// Point clouds: evaluation points on the host (eval_points.F90:129-225)
// ---------------------------------------------------------------------------
// The reference's candidate order: heapsort.F90 as it runs -- the initial
// heap covers elements 1..n-1 only, and a value moved through the swap
// temporary is rounded to single precision (real(kind=PR), PR = SP without
// -DDOUBLE_PRECISION).  plist/rlist are 1-based.
void heapsort_reference(int n, int *plist, double *rlist) {
  auto sift = [&](int root, int end) {
    for (;;) {
      int child = root * 2;
      if (child > end) break;
      if (child < end && rlist[child] < rlist[child + 1]) child++;
      if (!(rlist[root] < rlist[child])) break;
      const float t = (float)rlist[root];
      rlist[root] = rlist[child];
      rlist[child] = (double)t;
      std::swap(plist[root], plist[child]);
      root = child;
    }
  };
  for (int start = n / 2; start > 0; start--) sift(start, n - 1);
  for (int end = n; end > 1;) {
    const float t = (float)rlist[end];
    rlist[end] = rlist[1];
    rlist[1] = (double)t;
    std::swap(plist[end], plist[1]);
    end--;
    sift(1, end);
  }
}

struct OriginChains {
  std::vector<int> count;       // [nrays]
  std::vector<int> proj;        // ray-major
  std::vector<double> step;
};

// the evaluation points of one origin, all rays
void trace_origin(int p, int npts, const double *xyz, const int *etype, int nside, int nrays, const double *vec,
                  double theta_crit, int order, std::vector<int> &rb, std::vector<double> &ra,
                  std::vector<std::pair<double, int>> &srt, OriginChains &out) {
  const double *o = xyz + 3 * (size_t)p;
  const int n1 = npts - 1;
  int kk = 0;
  for (int i = 0; i < npts; i++) {
    if (i == p) continue;
    kk++;
    const double r0 = xyz[3 * (size_t)i] - o[0], r1 = xyz[3 * (size_t)i + 1] - o[1], r2 = xyz[3 * (size_t)i + 2] - o[2];
    ra[kk] = std::sqrt(r0 * r0 + r1 * r1 + r2 * r2);
    rb[kk] = i;
  }
  if (order == PDRCHEM_ORDER_REFERENCE) {
    heapsort_reference(n1, rb.data(), ra.data());
  } else {
    for (int k = 1; k <= n1; k++) srt[k - 1] = {ra[k], rb[k]};
    std::sort(srt.begin(), srt.begin() + n1);
    for (int k = 1; k <= n1; k++) {
      ra[k] = srt[k - 1].first;
      rb[k] = srt[k - 1].second;
    }
  }
  const double radius = n1 > 0 ? ra[n1] : 0.0;
  std::vector<double> ep(3 * nrays, 0.0), prev(3 * nrays);
  std::vector<char> kill(nrays, 0);
  std::vector<std::vector<int>> pj(nrays);
  std::vector<std::vector<double>> st(nrays);
  for (int j = 0; j < nrays; j++) {
    prev[3 * j] = o[0];
    prev[3 * j + 1] = o[1];
    prev[3 * j + 2] = o[2];
  }
  for (int k = 1; k <= n1; k++) {
    const int id = rb[k];
    const double rv[3] = {xyz[3 * (size_t)id] - o[0], xyz[3 * (size_t)id + 1] - o[1], xyz[3 * (size_t)id + 2] - o[2]};
    const int j = pix_of(nside, rv);
    if (kill[j]) continue;
    double *e = &ep[3 * j];
    double h[3];
    for (int c = 0; c < 3; c++) h[c] = 1.1 * radius * vec[3 * j + c];
    const double a0 = rv[0] - e[0], a1 = rv[1] - e[1], a2 = rv[2] - e[2];
    const double b0 = h[0] - e[0], b1 = h[1] - e[1], b2 = h[2] - e[2];
    const double dot = a0 * b0 + a1 * b1 + a2 * b2;
    const double na = std::sqrt(a0 * a0 + a1 * a1 + a2 * a2), nb = std::sqrt(b0 * b0 + b1 * b1 + b2 * b2);
    const double ang = std::acos(dot / (na * nb));
    if (!(ang <= theta_crit)) continue;
    if (h[2] != 0.0) {
      e[2] = (h[0] * h[2] * rv[0] + h[1] * h[2] * rv[1] + (h[2] * h[2]) * rv[2]) /
             (h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
      e[1] = e[2] * h[1] / h[2];
      e[0] = e[2] * h[0] / h[2];
    } else if (h[0] == 0.0) {
      e[0] = 0.0;
      e[1] = rv[1];
      e[2] = 0.0;
    } else if (h[1] == 0.0) {
      e[0] = rv[0];
      e[1] = 0.0;
      e[2] = 0.0;
    } else {
      e[2] = 0.0;
      e[0] = ((h[0] * h[0]) * rv[0] + h[0] * h[1] * rv[1]) / (h[0] * h[0] + h[1] * h[1]);
      e[1] = e[0] * h[1] / h[0];
    }
    const double c0 = e[0] + o[0], c1 = e[1] + o[1], c2 = e[2] + o[2];
    double *q = &prev[3 * j];
    pj[j].push_back(id);
    st[j].push_back(std::sqrt((q[0] - c0) * (q[0] - c0) + (q[1] - c1) * (q[1] - c1) + (q[2] - c2) * (q[2] - c2)));
    q[0] = c0;
    q[1] = c1;
    q[2] = c2;
    if (etype && etype[id] == 2) kill[j] = 1;
  }
  out.count.assign(nrays, 0);
  out.proj.clear();
  out.step.clear();
  for (int j = 0; j < nrays; j++) {
    out.count[j] = (int)pj[j].size();
    out.proj.insert(out.proj.end(), pj[j].begin(), pj[j].end());
    out.step.insert(out.step.end(), st[j].begin(), st[j].end());
  }
}

// every origin, threads over origins; CSR in (origin, ray) order
void trace_points_host(int npts, const double *xyz, const int *etype, int level, double theta_crit, int order,
                       std::vector<int> &count, std::vector<long long> &offs, std::vector<int> &proj,
                       std::vector<double> &step) {
  const int nside = 1 << level, nrays = 12 * nside * nside;
  std::vector<double> vec(3 * nrays);
  for (int j = 0; j < nrays; j++) pix_centre(nside, j, &vec[3 * j]);
  std::vector<OriginChains> res(npts);
  unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::atomic<int> next(0);
  auto work = [&]() {
    std::vector<int> rb(npts + 1);
    std::vector<double> ra(npts + 1);
    std::vector<std::pair<double, int>> srt(npts + 1);
    for (int p; (p = next.fetch_add(1)) < npts;)
      trace_origin(p, npts, xyz, etype, nside, nrays, vec.data(), theta_crit, order, rb, ra, srt, res[p]);
  };
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nth; t++) th.emplace_back(work);
  work();
  for (auto &t : th) t.join();
  count.assign((size_t)npts * nrays, 0);
  offs.assign((size_t)npts * nrays + 1, 0);
  long long tot = 0;
  for (int p = 0; p < npts; p++)
    for (int j = 0; j < nrays; j++) {
      offs[(size_t)p * nrays + j] = tot;
      count[(size_t)p * nrays + j] = res[p].count[j];
      tot += res[p].count[j];
    }
  offs[(size_t)npts * nrays] = tot;
  proj.resize(tot);
  step.resize(tot);
  for (int p = 0; p < npts; p++) {
    std::copy(res[p].proj.begin(), res[p].proj.end(), proj.begin() + offs[(size_t)p * nrays]);
    std::copy(res[p].step.begin(), res[p].step.end(), step.begin() + offs[(size_t)p * nrays]);
  }
}

// ---------------------------------------------------------------------------
// device kernels
// ---------------------------------------------------------------------------
struct RayGrid {
  int nx, ny, nz, nrays;
  double h, theta_crit;
};

// Walk ray j of every origin.  tab[k]: offset packed as (ox+128) |
// (oy+128)<<8 | (oz+128)<<16; r2[k] = |o|^2; entries of ray j are
// tab[ptr[j] .. ptr[j+1]), sorted by (|o|^2, linear offset).
template <bool FILL>
__global__ void __launch_bounds__(256) ray_walk_kernel(RayGrid g, const unsigned *__restrict__ tab,
                                                       const int *__restrict__ r2, const int *__restrict__ ptr,
                                                       const double *__restrict__ vec, int *__restrict__ count,
                                                       const long long *__restrict__ offs, int *__restrict__ proj,
                                                       double *__restrict__ step) {
  const int j = blockIdx.y;
  const int npts = g.nx * g.ny * g.nz;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const int px = p % g.nx, py = (p / g.nx) % g.ny, pz = p / (g.nx * g.ny);
  const double ox0 = px * g.h, oy0 = py * g.h, oz0 = pz * g.h;
  // distance to the farthest point: a corner (the last distance of the sort)
  double radius = 0.0;
  int r2max = 0;
  for (int c = 0; c < 8; c++) {
    const int dx = ((c & 1) ? g.nx - 1 : 0) - px, dy = ((c & 2) ? g.ny - 1 : 0) - py,
              dz = ((c & 4) ? g.nz - 1 : 0) - pz;
    const double rx = dx * g.h, ry = dy * g.h, rz = dz * g.h;
    radius = fmax(radius, sqrt(rx * rx + ry * ry + rz * rz));
    r2max = max(r2max, dx * dx + dy * dy + dz * dz);
  }
  const double hv0 = 1.1 * radius * vec[3 * j], hv1 = 1.1 * radius * vec[3 * j + 1],
               hv2 = 1.1 * radius * vec[3 * j + 2];
  double e0 = 0.0, e1 = 0.0, e2 = 0.0;        // evaluation point, relative to the origin
  double q0 = ox0, q1 = oy0, q2 = oz0;        // previous evaluation point, absolute
  int n = 0;
  long long w = FILL ? offs[(long long)p * g.nrays + j] : 0;
  const int kend = ptr[j + 1];
  for (int k = ptr[j]; k < kend; k++) {
    if (r2[k] > r2max) break;
    const unsigned e = tab[k];
    const int dx = (int)(e & 255u) - 128, dy = (int)((e >> 8) & 255u) - 128, dz = (int)((e >> 16) & 255u) - 128;
    const int tx = px + dx, ty = py + dy, tz = pz + dz;
    if (tx < 0 || tx >= g.nx || ty < 0 || ty >= g.ny || tz < 0 || tz >= g.nz) continue;
    const double rv0 = dx * g.h, rv1 = dy * g.h, rv2 = dz * g.h;
    const double a0 = rv0 - e0, a1 = rv1 - e1, a2 = rv2 - e2;
    const double b0 = hv0 - e0, b1 = hv1 - e1, b2 = hv2 - e2;
    const double dot = a0 * b0 + a1 * b1 + a2 * b2;
    const double na = sqrt(a0 * a0 + a1 * a1 + a2 * a2), nb = sqrt(b0 * b0 + b1 * b1 + b2 * b2);
    const double ang = acos(dot / (na * nb));
    if (!(ang <= g.theta_crit)) continue;
    if (hv2 != 0.0) {
      e2 = (hv0 * hv2 * rv0 + hv1 * hv2 * rv1 + (hv2 * hv2) * rv2) / (hv0 * hv0 + hv1 * hv1 + hv2 * hv2);
      e1 = e2 * hv1 / hv2;
      e0 = e2 * hv0 / hv2;
    } else if (hv0 == 0.0) {
      e0 = 0.0;
      e1 = rv1;
      e2 = 0.0;
    } else if (hv1 == 0.0) {
      e0 = rv0;
      e1 = 0.0;
      e2 = 0.0;
    } else {
      e2 = 0.0;
      e0 = ((hv0 * hv0) * rv0 + hv0 * hv1 * rv1) / (hv0 * hv0 + hv1 * hv1);
      e1 = e0 * hv1 / hv0;
    }
    n++;
    if (FILL) {
      const double c0 = e0 + ox0, c1 = e1 + oy0, c2 = e2 + oz0;
      proj[w] = tx + g.nx * (ty + g.ny * tz);
      step[w] = sqrt((q0 - c0) * (q0 - c0) + (q1 - c1) * (q1 - c1) + (q2 - c2) * (q2 - c2));
      q0 = c0;
      q1 = c1;
      q2 = c2;
      w++;
    }
  }
  if (!FILL) count[(long long)p * g.nrays + j] = n;
}

struct ColumnArgs {
  int npts, nrays, nspec;
  int sel[5];          // abundance slots of H2, HD, CO, C, S (-1: absent)
  double av_fac;
  const long long *offs;
  const int *count, *proj;
  const double *step, *rho, *ab;
  double *Av, *cols;   // [npts][nrays], [npts][nrays][5] (cols may be null)
};

__global__ void __launch_bounds__(256) ray_columns_kernel(ColumnArgs a) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)a.npts * a.nrays) return;
  const int p = (int)(t / a.nrays);
  const long long o = a.offs[t];
  const int n = a.count[t];
  double cd = 0.0, c[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  int q0 = p;
  for (int i = 0; i < n; i++) {
    const int q1 = a.proj[o + i];
    const double s = a.step[o + i];
    const double r0 = a.rho[q0], r1 = a.rho[q1];
    cd = cd + ((r0 + r1) / 2.) * s * kPc;
    if (a.cols) {
#pragma unroll
      for (int k = 0; k < 5; k++) {
        if (a.sel[k] < 0) continue;
        const double x0 = a.ab[(long long)q0 * a.nspec + a.sel[k]], x1 = a.ab[(long long)q1 * a.nspec + a.sel[k]];
        c[k] = c[k] + s * kPc * (r0 * x0 + r1 * x1) / 2.;
      }
    }
    q0 = q1;
  }
  a.Av[t] = cd * a.av_fac;
  if (a.cols)
    for (int k = 0; k < 5; k++) a.cols[t * 5 + k] = c[k];
}

// per-ray surface field, attenuated field and GUESS_TEMP temperature
// (3DPDR.F90:1667-1700); gray[j] is the surface field of ray j
__global__ void uvfield_kernel(int npts, int nrays, const double *__restrict__ gray, const double *__restrict__ Av,
                               double uv_fac, double *__restrict__ G, double *__restrict__ uv,
                               double *__restrict__ T) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  double u = 0.0;
  for (int j = 0; j < nrays; j++) {
    const double g = gray[j];
    G[(long long)p * nrays + j] = g;
    u = u + g * exp(-Av[(long long)p * nrays + j] * uv_fac);
    if (u < 1.0e-50) u = 0.0;
  }
  uv[p] = u;
  T[p] = 10.0 * (1.0 + pow(2. * u, 1.0 / 3.0));
}

__global__ void fill_kernel(double *x, long long n, double v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] = v;
}
__global__ void zero_u32_kernel(unsigned *x) { *x = 0u; }
__global__ void fill_u8_kernel(unsigned char *x, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0;
}
__global__ void zero_u64_kernel(unsigned long long *x, int n) {
  if ((int)threadIdx.x < n) x[threadIdx.x] = 0ull;
}

// rows of src selected by idx -> dst (gather) or back (scatter)
template <bool SCATTER>
__global__ void rows_kernel(double *__restrict__ dst, const double *__restrict__ src, const int *__restrict__ idx,
                            int nrows, int rowlen) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nrows * rowlen) return;
  const int r = (int)(t / rowlen), c = (int)(t % rowlen);
  const long long q = (long long)idx[r] * rowlen + c;
  if (SCATTER)
    dst[q] = src[t];
  else
    dst[t] = src[q];
}

// which points to solve: those not yet converged (a converged point keeps
// its abundances, as the reference keeps them, 3DPDR.F90:628)
__global__ void activity_kernel(int npts, const unsigned char *__restrict__ conv, unsigned char *__restrict__ active) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  active[p] = conv[p] ? 0 : 1;
}

// after a solve of the active points (compact index c -> point idx[c]):
// convergence test against the previous abundances and write-back
__global__ void retire_kernel(int nact, int nspec, const int *__restrict__ idx, const double *__restrict__ ab_c,
                              const int *__restrict__ status_c, double *__restrict__ ab,
                              unsigned char *__restrict__ conv, double tol,
                              double floor_abund, unsigned *__restrict__ nconv, unsigned *__restrict__ nfail,
                              const int *__restrict__ stats_c, unsigned long long *__restrict__ ctr) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nact) return;
  for (int k = 0; k < 5; k++) atomicAdd(&ctr[k], (unsigned long long)stats_c[(long long)c * 8 + k]);
  const int p = idx[c];
  double dmax = 0.0;
  for (int i = 0; i < nspec; i++) {
    const double xn = ab_c[(long long)c * nspec + i], xo = ab[(long long)p * nspec + i];
    const double m = fmax(fabs(xn), fabs(xo));
    if (m > floor_abund) dmax = fmax(dmax, fabs(xn - xo) / m);
    ab[(long long)p * nspec + i] = xn;
  }
  const bool cv = status_c[c] == 0 && dmax < tol;
  conv[p] = cv ? 1 : 0;
  if (cv) atomicAdd(nconv, 1u);
  if (status_c[c] != 0) atomicAdd(nfail, 1u);
}

struct ToLong {
  __host__ __device__ long long operator()(int v) const { return v; }
};

inline unsigned blocks_for(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// ---------------------------------------------------------------------------
// thermal-balance state of a cloud (pdrchem_cloud_thermal_*)
// ---------------------------------------------------------------------------
struct CloudThermal {
  pdrchem_coolant_tables *tables = nullptr;
  pdrchem::LvgTables tab{};
  pdrchem_thermal_params prm{};
  int heat_index[13] = {};
  int nlevtot = 0;
  double *d_pop = nullptr, *d_sol = nullptr, *d_S = nullptr;  // [npts][nlevtot], S [npts][nrays][nlevtot]
  double *d_cool = nullptr, *d_total = nullptr;               // [npts][4], [npts]
  double *d_lines = nullptr, *d_heat = nullptr;               // [npts][17], [npts][12]
  double *d_tau = nullptr;                                    // [npts][17] line optical depths
  double *d_fmean = nullptr, *d_fratio = nullptr;             // [npts]
  double *c_heat = nullptr, *c_uv = nullptr;                  // compact heating terms, field
  pdrchem::ThermalPoint *d_state = nullptr;
  unsigned char *d_lc = nullptr, *d_cconv = nullptr, *d_skip = nullptr;
  unsigned *d_counts = nullptr;                               // level-converged, per coolant x4, thermally converged
  // the reference's loop variables (3DPDR.F90:584-600)
  int iteration = 0, levpop_iteration = 0;
  bool level_conv = false, first_time = true, started = false, done = false;
  int pct[4] = {0, 0, 0, 0};                                  // CII, CI, OI, CO percentages as integers
  ~CloudThermal() { pdrchem_coolant_tables_destroy(tables); }
};

// ---------------------------------------------------------------------------
// the cloud object
// ---------------------------------------------------------------------------
struct pdrchem_cloud {
  int net, nspec, nreac, nelect, nx, ny, nz, npts, nrays, level;
  double h, theta_crit;
  pdrchem_phys phys;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // ray tracer
  int *d_count = nullptr, *d_proj = nullptr;
  long long *d_offs = nullptr, total = 0;
  double *d_step = nullptr;
  // per point
  double *d_rho = nullptr, *d_T = nullptr, *d_Td = nullptr, *d_uv = nullptr;
  double *d_G = nullptr, *d_Av = nullptr, *d_cols = nullptr, *d_ab = nullptr;
  unsigned char *d_conv = nullptr, *d_active = nullptr;
  int *d_idx = nullptr, *d_nsel = nullptr;
  // compact buffers of the active points
  double *c_T = nullptr, *c_Td = nullptr, *c_G = nullptr, *c_Av = nullptr, *c_cols = nullptr, *c_rho = nullptr;
  double *c_ab = nullptr, *c_rate = nullptr;
  double *d_zeta = nullptr, *c_zeta = nullptr;  // per-point cosmic-ray rate (optional)
  double *d_zeta_store = nullptr;
  int *c_status = nullptr, *c_stats = nullptr, *d_neg = nullptr;
  unsigned long long *d_ctr = nullptr;
  unsigned *d_counter = nullptr, *d_nconv = nullptr, *d_nfail = nullptr;
  void *d_cub = nullptr;
  size_t cub_bytes = 0;
  int sel[5] = {-1, -1, -1, -1, -1};
  double av_fac = 0.0, uv_fac = 0.0;
  bool have_ab = false;
  pdrchem_server *server = nullptr;  // when set, solves go through this request loop
  CloudThermal *th = nullptr;        // thermal balance, when initialised
  std::string err;
  std::vector<void *> allocs;

  template <class T>
  bool alloc(T **p, size_t n, const char *what) {
    void *q = nullptr;
    if (!ok(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)), what)) return false;
    allocs.push_back(q);
    *p = (T *)q;
    return true;
  }
  ~pdrchem_cloud() {
    // teardown: nothing to report to, so the return codes are dropped
    if (stream) (void)hipStreamSynchronize(stream);
    delete th;
    for (void *q : allocs) (void)hipFree(q);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
  bool fail(const char *what) {
    err = t_err.empty() ? std::string(what) : std::string(what) + ": " + t_err;
    return false;
  }
};

namespace {

bool build_rays(pdrchem_cloud *c) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, nside = 1 << c->level, nrays = c->nrays;
  // offsets of every pixel, sorted by (|o|^2, linear offset)
  struct Ent {
    int r2;
    long long lin;
    unsigned packed;
  };
  std::vector<std::vector<Ent>> per(nrays);
  for (int dz = -(nz - 1); dz <= nz - 1; dz++)
    for (int dy = -(ny - 1); dy <= ny - 1; dy++)
      for (int dx = -(nx - 1); dx <= nx - 1; dx++) {
        if (!dx && !dy && !dz) continue;
        const double r[3] = {dx * c->h, dy * c->h, dz * c->h};
        const int j = pix_of(nside, r);
        per[j].push_back(Ent{dx * dx + dy * dy + dz * dz, ((long long)dz * ny + dy) * nx + dx,
                             (unsigned)(dx + 128) | ((unsigned)(dy + 128) << 8) | ((unsigned)(dz + 128) << 16)});
      }
  std::vector<unsigned> tab;
  std::vector<int> r2, ptr(nrays + 1, 0);
  for (int j = 0; j < nrays; j++) {
    std::sort(per[j].begin(), per[j].end(),
              [](const Ent &a, const Ent &b) { return a.r2 != b.r2 ? a.r2 < b.r2 : a.lin < b.lin; });
    ptr[j] = (int)tab.size();
    for (const Ent &e : per[j]) {
      tab.push_back(e.packed);
      r2.push_back(e.r2);
    }
  }
  ptr[nrays] = (int)tab.size();
  std::vector<double> vec(3 * nrays);
  for (int j = 0; j < nrays; j++) pix_centre(nside, j, &vec[3 * j]);

  unsigned *d_tab = nullptr;
  int *d_r2 = nullptr, *d_ptr = nullptr;
  double *d_vec = nullptr;
  const long long nt = (long long)c->npts * nrays;
  if (!c->alloc(&d_tab, tab.size(), "ray table") || !c->alloc(&d_r2, r2.size(), "ray table") ||
      !c->alloc(&d_ptr, ptr.size(), "ray table") || !c->alloc(&d_vec, vec.size(), "ray vectors") ||
      !c->alloc(&c->d_count, nt, "ray counts") || !c->alloc(&c->d_offs, nt + 1, "ray offsets"))
    return c->fail("allocation");
  hipStream_t s = c->stream;
  if (!ok(hipMemcpyAsync(d_tab, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, s), "copy") ||
      !ok(hipMemcpyAsync(d_r2, r2.data(), r2.size() * 4, hipMemcpyHostToDevice, s), "copy") ||
      !ok(hipMemcpyAsync(d_ptr, ptr.data(), ptr.size() * 4, hipMemcpyHostToDevice, s), "copy") ||
      !ok(hipMemcpyAsync(d_vec, vec.data(), vec.size() * 8, hipMemcpyHostToDevice, s), "copy"))
    return c->fail("upload of ray tables");
  RayGrid g{nx, ny, nz, nrays, c->h, c->theta_crit};
  const dim3 grid(blocks_for(c->npts, 256), nrays);
  hipLaunchKernelGGL(ray_walk_kernel<false>, grid, dim3(256), 0, s, g, d_tab, d_r2, d_ptr, d_vec, c->d_count,
                     nullptr, nullptr, nullptr);
  if (!ok(hipGetLastError(), "ray count kernel")) return c->fail("ray tracer");
  // exclusive scan of counts into 64-bit offsets (offs[nt] = total)
  size_t tb = 0;
  hipcub::TransformInputIterator<long long, ToLong, const int *> in(c->d_count, ToLong());
  if (!ok(hipcub::DeviceScan::InclusiveSum(nullptr, tb, in, c->d_offs + 1, nt, s), "scan size"))
    return c->fail("ray tracer");
  void *tmp = nullptr;
  if (!c->alloc((char **)&tmp, tb, "scan workspace")) return c->fail("allocation");
  if (!ok(hipMemsetAsync(c->d_offs, 0, sizeof(long long), s), "memset") ||
      !ok(hipcub::DeviceScan::InclusiveSum(tmp, tb, in, c->d_offs + 1, nt, s), "scan"))
    return c->fail("ray tracer");
  if (!ok(hipMemcpyAsync(&c->total, c->d_offs + nt, sizeof(long long), hipMemcpyDeviceToHost, s), "copy") ||
      !ok(hipStreamSynchronize(s), "sync"))
    return c->fail("ray tracer");
  if (!c->alloc(&c->d_proj, c->total, "evaluation points") || !c->alloc(&c->d_step, c->total, "evaluation points"))
    return c->fail("allocation");
  hipLaunchKernelGGL(ray_walk_kernel<true>, grid, dim3(256), 0, s, g, d_tab, d_r2, d_ptr, d_vec, c->d_count,
                     c->d_offs, c->d_proj, c->d_step);
  if (!ok(hipGetLastError(), "ray fill kernel")) return c->fail("ray tracer");
  return true;
}

// evaluation points of a point cloud traced on the host, uploaded as CSR
bool upload_chains(pdrchem_cloud *c, const double *xyz, const int *etype, int order) {
  std::vector<int> cnt, pj;
  std::vector<long long> of;
  std::vector<double> st;
  trace_points_host(c->npts, xyz, etype, c->level, c->theta_crit, order, cnt, of, pj, st);
  c->total = (long long)pj.size();
  hipStream_t s = c->stream;
  if (!c->alloc(&c->d_count, cnt.size(), "ray counts") || !c->alloc(&c->d_offs, of.size(), "ray offsets") ||
      !c->alloc(&c->d_proj, pj.size(), "evaluation points") || !c->alloc(&c->d_step, st.size(), "evaluation points"))
    return c->fail("allocation");
  const bool good = ok(hipMemcpyAsync(c->d_count, cnt.data(), cnt.size() * 4, hipMemcpyHostToDevice, s), "copy") &&
                    ok(hipMemcpyAsync(c->d_offs, of.data(), of.size() * 8, hipMemcpyHostToDevice, s), "copy") &&
                    (pj.empty() || ok(hipMemcpyAsync(c->d_proj, pj.data(), pj.size() * 4, hipMemcpyHostToDevice, s),
                                      "copy")) &&
                    (st.empty() || ok(hipMemcpyAsync(c->d_step, st.data(), st.size() * 8, hipMemcpyHostToDevice, s),
                                      "copy")) &&
                    ok(hipStreamSynchronize(s), "sync");
  return good || c->fail("upload of evaluation points");
}

bool columns(pdrchem_cloud *c, bool with_species) {
  ColumnArgs a{};
  a.npts = c->npts;
  a.nrays = c->nrays;
  a.nspec = c->nspec;
  for (int k = 0; k < 5; k++) a.sel[k] = c->sel[k];
  a.av_fac = c->av_fac;
  a.offs = c->d_offs;
  a.count = c->d_count;
  a.proj = c->d_proj;
  a.step = c->d_step;
  a.rho = c->d_rho;
  a.ab = c->d_ab;
  a.Av = c->d_Av;
  a.cols = with_species ? c->d_cols : nullptr;
  hipLaunchKernelGGL(ray_columns_kernel, dim3(blocks_for((long long)c->npts * c->nrays, 256)), dim3(256), 0,
                     c->stream, a);
  return ok(hipGetLastError(), "column kernel") || c->fail("columns");
}

}  // namespace

extern "C" {

static pdrchem_cloud *create_common(int net, int nx, int ny, int nz, double spacing_pc, int npts_pc,
                                    const double *xyz, const int *etype, int order, const double *density, int level,
                                    double theta_crit, int fieldchoice, const double *gext, double dust_temperature,
                                    const pdrchem_phys *phys, const int *shield_slots, int nelect) {
  t_err.clear();
  int nspec = 0, nreac = 0;
  const int shapes[3][2] = {{33, 329}, {215, 2926}, {58, 634}};
  if (net < 0 || net > 2 || pdrchem_network_id(shapes[net][0], shapes[net][1]) != net) {
    t_err = "unsupported network id";
    return nullptr;
  }
  nspec = shapes[net][0];
  nreac = shapes[net][1];
  const bool grid = xyz == nullptr;
  if ((grid && (nx < 2 || ny < 2 || nz < 2 || nx > 128 || ny > 128 || nz > 128 || !(spacing_pc > 0))) ||
      (!grid && (npts_pc < 2 || npts_pc > 200000 ||
                 (order != PDRCHEM_ORDER_REFERENCE && order != PDRCHEM_ORDER_DISTANCE))) ||
      !density || !gext || !phys || level < 0 || level > 3 || !(theta_crit > 0) ||
      theta_crit >= 1.5707963267948966 || nelect < 1 || nelect > nspec || (fieldchoice != 0 && fieldchoice != 1)) {
    t_err = "bad argument to pdrchem_cloud_create (grid sides 2..128 or 2..200000 points, HEALPix level 0..3, "
            "0 < theta_crit < pi/2)";
    return nullptr;
  }
  pdrchem_cloud *c = new pdrchem_cloud();
  c->net = net;
  c->nspec = nspec;
  c->nreac = nreac;
  c->nelect = nelect;
  c->nx = nx;
  c->ny = ny;
  c->nz = nz;
  c->npts = grid ? nx * ny * nz : npts_pc;
  c->level = level;
  c->nrays = 12 << (2 * level);
  c->h = spacing_pc;
  c->theta_crit = theta_crit;
  c->phys = *phys;
  c->av_fac = (double)6.289e-22f * phys->metallicity;  // input_parameters.F90:57 (single-precision literal)
  c->uv_fac = (double)3.02f;                           // input_parameters.F90:58
  for (int k = 0; k < 5; k++) c->sel[k] = shield_slots ? shield_slots[k] : -1;
  const int n = c->npts, nr = c->nrays;
  const long long nc = (long long)n * nr * 5;
  bool good = ok(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream") &&
              ok(hipEventCreate(&c->ev0), "event") && ok(hipEventCreate(&c->ev1), "event") &&
              c->alloc(&c->d_rho, n, "density") && c->alloc(&c->d_T, n, "T") && c->alloc(&c->d_Td, n, "Td") &&
              c->alloc(&c->d_uv, n, "uv") && c->alloc(&c->d_G, (size_t)n * nr, "G") &&
              c->alloc(&c->d_Av, (size_t)n * nr, "Av") && c->alloc(&c->d_cols, nc, "columns") &&
              c->alloc(&c->d_ab, (size_t)n * nspec, "abundances") &&
              c->alloc(&c->d_conv, n, "flags") && c->alloc(&c->d_active, n, "flags") &&
              c->alloc(&c->d_idx, n, "index") && c->alloc(&c->d_nsel, 1, "count") &&
              c->alloc(&c->c_T, n, "compact") && c->alloc(&c->c_Td, n, "compact") &&
              c->alloc(&c->c_G, (size_t)n * nr, "compact") && c->alloc(&c->c_Av, (size_t)n * nr, "compact") &&
              c->alloc(&c->c_cols, nc, "compact") && c->alloc(&c->c_rho, n, "compact") &&
              c->alloc(&c->c_ab, (size_t)n * nspec, "compact") && c->alloc(&c->c_rate, (size_t)n * nreac, "rates") &&
              c->alloc(&c->c_status, n, "status") && c->alloc(&c->c_stats, (size_t)n * 8, "counters") &&
              c->alloc(&c->d_ctr, 8, "counters") && c->alloc(&c->d_neg, 1, "count") &&
              c->alloc(&c->d_counter, 1, "count") && c->alloc(&c->d_nconv, 1, "count") &&
              c->alloc(&c->d_nfail, 1, "count");
  if (!good) {
    delete c;
    return nullptr;
  }
  hipStream_t s = c->stream;
  // per-ray surface field (UNI: only the ray most opposed to gext)
  std::vector<double> vec(3 * nr), gray(nr, 0.0);
  for (int j = 0; j < nr; j++) pix_centre(1 << level, j, &vec[3 * j]);
  int jmin = 0;
  double best = 0.0;
  for (int j = 0; j < nr; j++) {
    const double d = -(gext[0] * vec[3 * j] + gext[1] * vec[3 * j + 1] + gext[2] * vec[3 * j + 2]);
    if (d >= best) {
      best = d;
      jmin = j;
    }
  }
  for (int j = 0; j < nr; j++) {
    if (fieldchoice == 0)
      gray[j] = gext[0] / (double)nr;
    else if (j == jmin)
      gray[j] = -(gext[0] * vec[3 * j] + gext[1] * vec[3 * j + 1] + gext[2] * vec[3 * j + 2]);
  }
  double *d_gray = nullptr;
  good = c->alloc(&d_gray, nr, "field") &&
         ok(hipMemcpyAsync(d_gray, gray.data(), nr * 8, hipMemcpyHostToDevice, s), "copy") &&
         ok(hipMemcpyAsync(c->d_rho, density, (size_t)n * 8, hipMemcpyHostToDevice, s), "copy") &&
         ok(hipMemsetAsync(c->d_conv, 0, n, s), "memset") &&
         (grid ? build_rays(c) : upload_chains(c, xyz, etype, order)) && columns(c, false);
  if (good) {
    hipLaunchKernelGGL(uvfield_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, nr, d_gray, c->d_Av,
                       c->uv_fac, c->d_G, c->d_uv, c->d_T);
    hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, c->d_Td, (long long)n,
                       dust_temperature);
    good = ok(hipGetLastError(), "field kernels") && ok(hipStreamSynchronize(s), "sync");
  }
  if (good) {
    // stream-compaction workspace
    hipcub::CountingInputIterator<int> it(0);
    if (!ok(hipcub::DeviceSelect::Flagged(nullptr, c->cub_bytes, it, c->d_active, c->d_idx, c->d_nsel, n, s),
            "select size") ||
        !c->alloc((char **)&c->d_cub, c->cub_bytes, "select workspace"))
      good = false;
  }
  if (!good) {
    t_err = c->err.empty() ? t_err : c->err;
    delete c;
    return nullptr;
  }
  return c;
}

pdrchem_cloud *pdrchem_cloud_create(int net, int nx, int ny, int nz, double spacing_pc, const double *density,
                                    int level, double theta_crit, int fieldchoice, const double *gext,
                                    double dust_temperature, const pdrchem_phys *phys, const int *shield_slots,
                                    int nelect) {
  return create_common(net, nx, ny, nz, spacing_pc, 0, nullptr, nullptr, 0, density, level, theta_crit, fieldchoice,
                       gext, dust_temperature, phys, shield_slots, nelect);
}

pdrchem_cloud *pdrchem_cloud_create_points(int net, int npts, const double *xyz, const int *etype, int order,
                                           const double *density, int level, double theta_crit, int fieldchoice,
                                           const double *gext, double dust_temperature, const pdrchem_phys *phys,
                                           const int *shield_slots, int nelect) {
  if (!xyz) {
    t_err = "pdrchem_cloud_create_points needs positions";
    return nullptr;
  }
  return create_common(net, 0, 0, 0, 0.0, npts, xyz, etype, order, density, level, theta_crit, fieldchoice, gext,
                       dust_temperature, phys, shield_slots, nelect);
}

long long pdrchem_trace_points(int npts, const double *xyz, const int *etype, int level, double theta_crit, int order,
                               int *count, long long *offsets, int *proj, double *step, long long capacity) {
  if (npts < 2 || !xyz || level < 0 || level > 3 || !(theta_crit > 0) || theta_crit >= 1.5707963267948966 ||
      (order != PDRCHEM_ORDER_REFERENCE && order != PDRCHEM_ORDER_DISTANCE))
    return -1;
  std::vector<int> cnt, pj;
  std::vector<long long> of;
  std::vector<double> st;
  trace_points_host(npts, xyz, etype, level, theta_crit, order, cnt, of, pj, st);
  if (count) std::copy(cnt.begin(), cnt.end(), count);
  if (offsets) std::copy(of.begin(), of.end(), offsets);
  const long long total = (long long)pj.size();
  if (capacity >= total) {
    if (proj) std::copy(pj.begin(), pj.end(), proj);
    if (step) std::copy(st.begin(), st.end(), step);
  }
  return total;
}

int pdrchem_cloud_set_abundances(pdrchem_cloud *c, const double *abundance) {
  if (!c || !abundance) return -1;
  c->err.clear();
  t_err.clear();
  const size_t b = (size_t)c->npts * c->nspec * 8;
  if (!ok(hipMemcpyAsync(c->d_ab, abundance, b, hipMemcpyHostToDevice, c->stream), "copy") ||
      !ok(hipMemsetAsync(c->d_conv, 0, c->npts, c->stream), "memset") || !columns(c, true) ||
      !ok(hipStreamSynchronize(c->stream), "sync")) {
    c->fail("set abundances");
    return -1;
  }
  c->have_ab = true;
  return 0;
}

int pdrchem_cloud_iterate(pdrchem_cloud *c, double t_end_years, double conv_tol, int precision,
                          pdrchem_cloud_stats *st) {
  if (!c) return -1;
  c->err.clear();
  t_err.clear();
  if (!c->have_ab) {
    c->err = "pdrchem_cloud_iterate before pdrchem_cloud_set_abundances";
    return -1;
  }
  if (!(t_end_years > 0) || !(conv_tol >= 0) || (precision != PDRCHEM_FP64 && precision != PDRCHEM_MIXED)) {
    c->err = "bad argument to pdrchem_cloud_iterate";
    return -1;
  }
  hipStream_t s = c->stream;
  const int n = c->npts, nr = c->nrays, ncol = c->nrays * 5;
  // 1. which points to solve
  hipLaunchKernelGGL(activity_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, c->d_conv, c->d_active);
  hipcub::CountingInputIterator<int> it(0);
  int nact = 0;
  if (!ok(hipGetLastError(), "activity kernel") ||
      !ok(hipcub::DeviceSelect::Flagged(c->d_cub, c->cub_bytes, it, c->d_active, c->d_idx, c->d_nsel, n, s),
          "select") ||
      !ok(hipMemcpyAsync(&nact, c->d_nsel, sizeof(int), hipMemcpyDeviceToHost, s), "copy") ||
      !ok(hipStreamSynchronize(s), "sync")) {
    c->fail("active set");
    return -1;
  }
  pdrchem_cloud_stats loc{};
  loc.npts = n;
  loc.solved = nact;
  float ms_rates = 0.f, ms_solve = 0.f;
  if (nact > 0) {
    // 2. gather the active points' inputs, rates on device
    hipLaunchKernelGGL(rows_kernel<false>, dim3(blocks_for(nact, 256)), dim3(256), 0, s, c->c_T, c->d_T, c->d_idx,
                       nact, 1);
    hipLaunchKernelGGL(rows_kernel<false>, dim3(blocks_for(nact, 256)), dim3(256), 0, s, c->c_Td, c->d_Td,
                       c->d_idx, nact, 1);
    hipLaunchKernelGGL(rows_kernel<false>, dim3(blocks_for(nact, 256)), dim3(256), 0, s, c->c_rho, c->d_rho,
                       c->d_idx, nact, 1);
    hipLaunchKernelGGL(rows_kernel<false>, dim3(blocks_for((long long)nact * nr, 256)), dim3(256), 0, s, c->c_G,
                       c->d_G, c->d_idx, nact, nr);
    hipLaunchKernelGGL(rows_kernel<false>, dim3(blocks_for((long long)nact * nr, 256)), dim3(256), 0, s, c->c_Av,
                       c->d_Av, c->d_idx, nact, nr);
    hipLaunchKernelGGL(rows_kernel<false>, dim3(blocks_for((long long)nact * ncol, 256)), dim3(256), 0, s,
                       c->c_cols, c->d_cols, c->d_idx, nact, ncol);
    hipLaunchKernelGGL(rows_kernel<false>, dim3(blocks_for((long long)nact * c->nspec, 256)), dim3(256), 0, s,
                       c->c_ab, c->d_ab, c->d_idx, nact, c->nspec);
    if (c->d_zeta)
      hipLaunchKernelGGL(rows_kernel<false>, dim3(blocks_for(nact, 256)), dim3(256), 0, s, c->c_zeta, c->d_zeta,
                         c->d_idx, nact, 1);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, (unsigned *)c->d_neg);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, c->d_counter);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, c->d_nconv);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, c->d_nfail);
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, c->d_ctr, 8);
    if (!ok(hipGetLastError(), "gather kernels") || !ok(hipEventRecord(c->ev0, s), "event")) {
      c->fail("gather");
      return -1;
    }
    if (pdrchem_rates_device_zeta(c->net, &c->phys, c->c_T, c->c_Td, nr, c->c_G, c->c_Av, c->c_cols,
                                  c->d_zeta ? c->c_zeta : nullptr, nact, c->c_rate, c->d_neg, s) != 0) {
      c->err = std::string("rates: ") + pdrchem_last_error();
      return -1;
    }
    if (!ok(hipEventRecord(c->ev1, s), "event")) {
      c->fail("rates");
      return -1;
    }
    hipEvent_t e2 = nullptr, e3 = nullptr;
    if (!ok(hipEventCreate(&e2), "event") || !ok(hipEventCreate(&e3), "event") ||
        !ok(hipEventRecord(e2, s), "event")) {
      c->fail("events");
      return -1;
    }
    const double rtol = chemistry_module_mp_relative_abundance_tolerance_;
    const double atol = chemistry_module_mp_absolute_abundance_tolerance_;
    double server_ms = -1.0;
    if (c->server) {
      // through the request loop: the compact batch as device requests of at
      // most one batch each, solved in place on the loop's streams
      const int chunk = pdrchem_server_max_batch(c->server);
      bool good = chunk > 0 && ok(hipStreamSynchronize(s), "sync");
      const auto t0 = std::chrono::steady_clock::now();
      for (int b = 0; good && b < nact; b += chunk) {
        const int m = std::min(chunk, nact - b);
        good = pdrchem_server_submit_device(c->server, c->net, m, c->c_ab + (size_t)b * c->nspec,
                                            c->c_rate + (size_t)b * c->nreac, c->c_rho + b, t_end_years,
                                            c->c_status + b, c->c_stats + (size_t)b * 8) >= 0;
      }
      good = good && pdrchem_server_progress(c->server, 1) >= 0;
      std::vector<int> tk(64), st(64);
      while (good && pdrchem_server_poll(c->server, 64, tk.data(), st.data(), nullptr) > 0) {
      }
      server_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      if (!good) {
        c->err = std::string("solve through the request loop: ") + pdrchem_server_last_error(c->server);
        (void)hipEventDestroy(e2);
        (void)hipEventDestroy(e3);
        return -1;
      }
    } else if (pdrchem_solve_device(c->net, c->c_ab, c->c_rate, c->c_rho, nact, rtol, atol, 0.0, t_end_years,
                                    c->nelect, c->c_status, c->c_stats, c->d_counter, precision, s) != 0) {
      c->err = std::string("solve: ") + pdrchem_last_error();
      (void)hipEventDestroy(e2);
      (void)hipEventDestroy(e3);
      return -1;
    }
    const bool recorded = ok(hipEventRecord(e3, s), "event");
    // 3. convergence, write-back
    hipLaunchKernelGGL(retire_kernel, dim3(blocks_for(nact, 256)), dim3(256), 0, s, nact, c->nspec, c->d_idx,
                       c->c_ab, c->c_status, c->d_ab, c->d_conv, conv_tol, 1.0e-20,
                       c->d_nconv, c->d_nfail, c->c_stats, c->d_ctr);
    unsigned cnt[3] = {0, 0, 0};
    unsigned long long ctr[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool good = recorded && ok(hipGetLastError(), "retire kernel") && columns(c, true) &&
                ok(hipMemcpyAsync(&cnt[0], c->d_nconv, 4, hipMemcpyDeviceToHost, s), "copy") &&
                ok(hipMemcpyAsync(&cnt[1], c->d_nfail, 4, hipMemcpyDeviceToHost, s), "copy") &&
                ok(hipMemcpyAsync(&cnt[2], c->d_neg, 4, hipMemcpyDeviceToHost, s), "copy") &&
                ok(hipMemcpyAsync(ctr, c->d_ctr, sizeof(ctr), hipMemcpyDeviceToHost, s), "copy") &&
                ok(hipStreamSynchronize(s), "sync");
    if (good) {
      // timings only: a failed query leaves -1 (not measured)
      if (hipEventElapsedTime(&ms_rates, c->ev0, c->ev1) != hipSuccess) ms_rates = -1.0f;
      if (hipEventElapsedTime(&ms_solve, e2, e3) != hipSuccess) ms_solve = -1.0f;
      if (server_ms >= 0) ms_solve = (float)server_ms;
    }
    (void)hipEventDestroy(e2);
    (void)hipEventDestroy(e3);
    if (!good) {
      c->fail("iteration");
      return -1;
    }
    loc.converged = (int)cnt[0];
    loc.failed = (int)cnt[1];
    loc.negative_rates = (int)cnt[2];
    loc.nst = (long long)ctr[0];
    loc.nfe = (long long)ctr[1];
    loc.nje = (long long)ctr[2];
    loc.nsetups = (long long)ctr[3];
    loc.nni = (long long)ctr[4];
  }
  loc.rates_ms = ms_rates;
  loc.solve_ms = ms_solve;
  if (st) *st = loc;
  return 0;
}

// thermal-balance fields of pdrchem_cloud_get
static int thermal_get(pdrchem_cloud *c, int field, double *out) {
  CloudThermal *th = c->th;
  if (!th) {
    c->err = "thermal field requested before pdrchem_cloud_thermal_init";
    return -1;
  }
  const size_t n = c->npts;
  hipStream_t s = c->stream;
  bool good = true;
  switch (field) {
    case PDRCHEM_CLOUD_HEATING:
      good = ok(hipMemcpyAsync(out, th->d_heat, n * 12 * 8, hipMemcpyDeviceToHost, s), "copy");
      break;
    case PDRCHEM_CLOUD_LINES:
      good = ok(hipMemcpyAsync(out, th->d_lines, n * pdrchem::LvgArgs::kLines * 8, hipMemcpyDeviceToHost, s), "copy");
      break;
    case PDRCHEM_CLOUD_OPTICAL_DEPTHS:
      good = ok(hipMemcpyAsync(out, th->d_tau, n * pdrchem::LvgArgs::kLines * 8, hipMemcpyDeviceToHost, s), "copy");
      break;
    case PDRCHEM_CLOUD_POPULATIONS:
      good = ok(hipMemcpyAsync(out, th->d_pop, n * th->nlevtot * 8, hipMemcpyDeviceToHost, s), "copy");
      break;
    case PDRCHEM_CLOUD_COOLING: {
      std::vector<double> cool(n * 4), tot(n);
      good = ok(hipMemcpyAsync(cool.data(), th->d_cool, n * 4 * 8, hipMemcpyDeviceToHost, s), "copy") &&
             ok(hipMemcpyAsync(tot.data(), th->d_total, n * 8, hipMemcpyDeviceToHost, s), "copy") &&
             ok(hipStreamSynchronize(s), "sync");
      for (size_t p = 0; good && p < n; p++) {
        for (int k = 0; k < 4; k++) out[p * 5 + k] = cool[p * 4 + k];
        out[p * 5 + 4] = tot[p];
      }
      break;
    }
    default: {
      std::vector<pdrchem::ThermalPoint> st(n);
      good = ok(hipMemcpyAsync(st.data(), th->d_state, n * sizeof(pdrchem::ThermalPoint), hipMemcpyDeviceToHost, s),
                "copy") &&
             ok(hipStreamSynchronize(s), "sync");
      for (size_t p = 0; good && p < n; p++)
        out[p] = field == PDRCHEM_CLOUD_PREV_TEMPERATURE ? st[p].Tprev : (double)st[p].converged;
    }
  }
  if (!good || !ok(hipStreamSynchronize(s), "sync")) {
    c->fail("get");
    return -1;
  }
  return 0;
}

int pdrchem_cloud_get(pdrchem_cloud *c, int field, double *out) {
  if (!c || !out) return -1;
  c->err.clear();
  const size_t n = c->npts, nr = c->nrays;
  const double *src = nullptr;
  size_t cnt = 0;
  switch (field) {
    case PDRCHEM_CLOUD_ABUNDANCE: src = c->d_ab; cnt = n * c->nspec; break;
    case PDRCHEM_CLOUD_AV: src = c->d_Av; cnt = n * nr; break;
    case PDRCHEM_CLOUD_G: src = c->d_G; cnt = n * nr; break;
    case PDRCHEM_CLOUD_COLUMNS: src = c->d_cols; cnt = n * nr * 5; break;
    case PDRCHEM_CLOUD_TEMPERATURE: src = c->d_T; cnt = n; break;
    case PDRCHEM_CLOUD_UVFIELD: src = c->d_uv; cnt = n; break;
    case PDRCHEM_CLOUD_DUST_TEMPERATURE: src = c->d_Td; cnt = n; break;
    case PDRCHEM_CLOUD_HEATING:
    case PDRCHEM_CLOUD_LINES:
    case PDRCHEM_CLOUD_POPULATIONS:
    case PDRCHEM_CLOUD_COOLING:
    case PDRCHEM_CLOUD_PREV_TEMPERATURE:
    case PDRCHEM_CLOUD_THERMAL_CONVERGED:
    case PDRCHEM_CLOUD_OPTICAL_DEPTHS:
      return thermal_get(c, field, out);
    default:
      c->err = "unknown field";
      return -1;
  }
  if (!ok(hipMemcpyAsync(out, src, cnt * 8, hipMemcpyDeviceToHost, c->stream), "copy") ||
      !ok(hipStreamSynchronize(c->stream), "sync")) {
    c->fail("get");
    return -1;
  }
  return 0;
}

int pdrchem_cloud_set_zeta(pdrchem_cloud *c, const double *zeta) {
  if (!c) return -1;
  c->err.clear();
  t_err.clear();
  if (!zeta) {
    c->d_zeta = nullptr;  // back to phys.zeta (the buffers stay allocated)
    return 0;
  }
  if (!c->c_zeta) {
    double *dz = nullptr, *cz = nullptr;
    if (!c->alloc(&dz, c->npts, "zeta") || !c->alloc(&cz, c->npts, "zeta")) {
      c->fail("allocation");
      return -1;
    }
    c->c_zeta = cz;
    c->d_zeta_store = dz;
  }
  if (!ok(hipMemcpyAsync(c->d_zeta_store, zeta, (size_t)c->npts * 8, hipMemcpyHostToDevice, c->stream), "copy") ||
      !ok(hipStreamSynchronize(c->stream), "sync")) {
    c->fail("set zeta");
    return -1;
  }
  c->d_zeta = c->d_zeta_store;
  hipLaunchKernelGGL(fill_u8_kernel, dim3(blocks_for(c->npts, 256)), dim3(256), 0, c->stream, c->d_conv, c->npts);
  return ok(hipGetLastError(), "flags") ? 0 : -1;
}

int pdrchem_cloud_set_server(pdrchem_cloud *c, pdrchem_server *server) {
  if (!c) return -1;
  c->server = server;
  return 0;
}

long long pdrchem_cloud_num_evaluation_points(const pdrchem_cloud *c) { return c ? c->total : -1; }

int pdrchem_cloud_evaluation_points(pdrchem_cloud *c, int *count, long long *offsets, int *proj, double *step) {
  if (!c) return -1;
  c->err.clear();
  hipStream_t s = c->stream;
  const size_t nt = (size_t)c->npts * c->nrays;
  bool good = (!count || ok(hipMemcpyAsync(count, c->d_count, nt * 4, hipMemcpyDeviceToHost, s), "copy")) &&
              (!offsets || ok(hipMemcpyAsync(offsets, c->d_offs, (nt + 1) * 8, hipMemcpyDeviceToHost, s), "copy")) &&
              (!proj || ok(hipMemcpyAsync(proj, c->d_proj, (size_t)c->total * 4, hipMemcpyDeviceToHost, s), "copy")) &&
              (!step || ok(hipMemcpyAsync(step, c->d_step, (size_t)c->total * 8, hipMemcpyDeviceToHost, s), "copy")) &&
              ok(hipStreamSynchronize(s), "sync");
  if (!good) {
    c->fail("evaluation points");
    return -1;
  }
  return 0;
}

int pdrchem_cloud_thermal_init(pdrchem_cloud *c, const pdrchem_coolant *coolants, const int *slots,
                               const int *heat_index, const pdrchem_thermal_params *prm) {
  using namespace pdrchem;
  if (!c) return -1;
  c->err.clear();
  t_err.clear();
  if (!coolants || !slots || !heat_index || !prm) {
    c->err = "bad argument to pdrchem_cloud_thermal_init (four coolants, 9 slots, 13 heating indices)";
    return -1;
  }
  for (int k = 0; k < 9; k++)
    if (slots[k] < 0 || slots[k] >= c->nspec) {
      c->err = "abundance slot out of range";
      return -1;
    }
  if (prm->pseudo_ray < -1 || prm->pseudo_ray >= c->nrays || !(prm->Tmin > 0) || !(prm->Tmax > prm->Tmin) || !(prm->Fcrit > 0)) {
    c->err = "bad thermal parameters";
    return -1;
  }
  delete c->th;
  c->th = new CloudThermal();
  CloudThermal *th = c->th;
  th->prm = *prm;
  for (int k = 0; k < 13; k++) th->heat_index[k] = heat_index[k];
  th->tables = pdrchem_coolant_tables_create(coolants, slots);
  if (!th->tables) {
    c->err = pdrchem_last_error();
    delete c->th;
    c->th = nullptr;
    return -1;
  }
  th->tab = *coolant_tables_dev(th->tables);
  const int off = th->tab.nlevtot;
  th->nlevtot = off;
  hipStream_t s = c->stream;
  const size_t n = c->npts;
  if (!c->alloc(&th->d_pop, n * off, "populations") || !c->alloc(&th->d_sol, n * off, "populations") ||
      !c->alloc(&th->d_S, n * c->nrays * off, "level columns") || !c->alloc(&th->d_cool, n * 4, "cooling") ||
      !c->alloc(&th->d_total, n, "cooling") || !c->alloc(&th->d_lines, n * LvgArgs::kLines, "lines") ||
      !c->alloc(&th->d_heat, n * 12, "heating") || !c->alloc(&th->d_tau, n * LvgArgs::kLines, "optical depths") ||
      !c->alloc(&th->d_fmean, n, "balance") ||
      !c->alloc(&th->d_fratio, n, "balance") || !c->alloc(&th->c_heat, n * 12, "heating") ||
      !c->alloc(&th->c_uv, n, "field") || !c->alloc(&th->d_state, n, "thermal state") ||
      !c->alloc(&th->d_lc, n, "flags") || !c->alloc(&th->d_cconv, n * 4, "flags") ||
      !c->alloc(&th->d_skip, n, "flags") || !c->alloc(&th->d_counts, 8, "counts") ||
      !ok(hipMemsetAsync(th->d_cool, 0, n * 4 * 8, s), "memset") ||
      !ok(hipMemsetAsync(th->d_total, 0, n * 8, s), "memset") ||
      !ok(hipMemsetAsync(th->d_lines, 0, n * LvgArgs::kLines * 8, s), "memset") ||
      !ok(hipMemsetAsync(th->d_heat, 0, n * 12 * 8, s), "memset") ||
      !ok(hipMemsetAsync(th->d_tau, 0, n * LvgArgs::kLines * 8, s), "memset") ||
      !ok(hipMemsetAsync(th->d_lc, 0, n, s), "memset")) {
    c->fail("thermal state");
    return -1;
  }
  // dust temperatures from the field (CALCULATE_DUST_TEMPERATURES), the
  // temperature bracket of GUESS_TEMP (3DPDR.F90:1702-1716)
  if ((prm->dust && launch_dust_temperature(c->npts, c->nrays, c->d_G, c->d_uv, c->d_Td, s) != 0) ||
      launch_init_thermal_state(c->npts, c->d_T, prm->Tmin, prm->Tmax, th->d_state, s) != 0 ||
      !ok(hipStreamSynchronize(s), "sync")) {
    c->fail("thermal set-up");
    return -1;
  }
  return 0;
}

// one ITERATION of the reference's outer loop (3DPDR.F90:590-1241)
int pdrchem_cloud_thermal_iterate(pdrchem_cloud *c, double t_end_years, int precision, pdrchem_thermal_stats *out) {
  using namespace pdrchem;
  if (!c) return -1;
  c->err.clear();
  t_err.clear();
  CloudThermal *th = c->th;
  if (!th || !c->have_ab) {
    c->err = "pdrchem_cloud_thermal_iterate needs pdrchem_cloud_thermal_init and abundances";
    return -1;
  }
  if (!(t_end_years > 0) || (precision != PDRCHEM_FP64 && precision != PDRCHEM_MIXED)) {
    c->err = "bad argument to pdrchem_cloud_thermal_iterate";
    return -1;
  }
  pdrchem_thermal_stats st{};
  if (th->done) {
    st.done = 1;
    st.iteration = th->iteration;
    if (out) *out = st;
    return 0;
  }
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = c->stream;
  const int n = c->npts;
  const LvgTables &t = th->tab;
  auto fail = [&](const char *what) {
    c->fail(what);
    return -1;
  };
  if (!th->started) {
    // LTE populations of every point before the first iteration (3DPDR.F90:520-552)
    if (launch_lte(t, n, c->d_T, c->d_ab, c->nspec, c->d_rho, nullptr, th->d_pop, s) != 0) return fail("LTE");
    th->started = true;
  }
  th->iteration++;
  th->levpop_iteration++;
  st.iteration = th->iteration;
  if (th->iteration > 1 && th->levpop_iteration == 1) {
    // three chemistry iterations of the points not thermally converged,
    // then LTE populations of those not converged in levels either
    // (3DPDR.F90:602-686)
    for (int k = 0; k < 3; k++) {
      if (launch_skip_flags(n, th->d_state, th->d_lc, 0, c->d_conv, s) != 0) return fail("flags");
      pdrchem_cloud_stats cs{};
      if (pdrchem_cloud_iterate(c, t_end_years, 0.0, precision, &cs) != 0) return -1;
      st.chemistry_solves += cs.solved;
      st.failed += cs.failed;
      st.nst += cs.nst;
      st.nfe += cs.nfe;
      st.nje += cs.nje;
      st.nsetups += cs.nsetups;
      st.nni += cs.nni;
    }
    if (launch_skip_flags(n, th->d_state, th->d_lc, 1, th->d_skip, s) != 0 ||
        launch_lte(t, n, c->d_T, c->d_ab, c->nspec, c->d_rho, th->d_skip, th->d_pop, s) != 0)
      return fail("LTE");
  } else {
    // level populations by escape probability (3DPDR.F90:688-850)
    const int p1 = th->prm.pseudo_ray;
    if (launch_skip_flags(n, th->d_state, th->d_lc, 1, th->d_skip, s) != 0 ||
        launch_level_columns(n, c->nrays, p1 >= 0 ? p1 : 0, p1 >= 0 ? 1 : c->nrays, c->d_offs, c->d_count,
                             c->d_proj, c->d_step, th->d_pop, th->nlevtot, th->d_S, s) != 0)
      return fail("level columns");
    LvgArgs a{};
    a.npts = n;
    a.nspec = c->nspec;
    a.nrays = c->nrays;
    a.pseudo_ray = p1;
    a.T = c->d_T;
    a.Td = c->d_Td;
    a.rho = c->d_rho;
    a.ab = c->d_ab;
    a.pop = th->d_pop;
    a.S = th->d_S;
    a.skip = th->d_skip;
    a.v_turb = c->phys.v_turb;
    a.metallicity = c->phys.metallicity;
    a.g2d = c->phys.g2d;
    a.co_fix = th->prm.co_fix;
    a.co_average_co = th->pct[0] == 100 && th->pct[1] == 100 && th->pct[2] == 100;
    a.levpop_iteration = th->levpop_iteration;
    a.solution = th->d_sol;
    a.cool = th->d_cool;
    a.lines = th->d_lines;
    a.tau_lines = th->d_tau;
    a.tau_ray = p1 >= 0 ? p1 : (c->nrays > 6 ? 6 : 0);
    for (int k = 0; k < 4; k++)
      if (launch_lvg(t, a, k, s) != 0) return fail("level populations");
    if (launch_total_cool(n, th->d_cool, th->d_total, s) != 0) return fail("cooling");
    st.lvg = 1;
  }
  // heating and the temperature update of the points not converged
  // (3DPDR.F90:852-1057)
  {
    if (launch_skip_flags(n, th->d_state, th->d_lc, 0, th->d_skip, s) != 0) return fail("flags");
    hipLaunchKernelGGL(activity_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, th->d_skip, c->d_active);
    hipcub::CountingInputIterator<int> it(0);
    int nact = 0;
    if (!ok(hipGetLastError(), "flags") ||
        !ok(hipcub::DeviceSelect::Flagged(c->d_cub, c->cub_bytes, it, c->d_active, c->d_idx, c->d_nsel, n, s),
            "select") ||
        !ok(hipMemcpyAsync(&nact, c->d_nsel, sizeof(int), hipMemcpyDeviceToHost, s), "copy") ||
        !ok(hipStreamSynchronize(s), "sync"))
      return fail("active set");
    st.thermal_points = nact;
    if (nact > 0) {
      const int nr = c->nrays, ncol = nr * 5;
      auto gather = [&](double *dst, const double *src, int w) {
        hipLaunchKernelGGL(rows_kernel<false>, dim3(blocks_for((long long)nact * w, 256)), dim3(256), 0, s, dst, src,
                           c->d_idx, nact, w);
      };
      gather(c->c_T, c->d_T, 1);
      gather(c->c_Td, c->d_Td, 1);
      gather(c->c_rho, c->d_rho, 1);
      gather(th->c_uv, c->d_uv, 1);
      gather(c->c_G, c->d_G, nr);
      gather(c->c_Av, c->d_Av, nr);
      gather(c->c_cols, c->d_cols, ncol);
      gather(c->c_ab, c->d_ab, c->nspec);
      if (c->d_zeta) gather(c->c_zeta, c->d_zeta, 1);
      hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, (unsigned *)c->d_neg);
      if (!ok(hipGetLastError(), "gather") ||
          pdrchem_rates_device_zeta(c->net, &c->phys, c->c_T, c->c_Td, nr, c->c_G, c->c_Av, c->c_cols,
                                    c->d_zeta ? c->c_zeta : nullptr, nact, c->c_rate, c->d_neg, s) != 0)
        return fail("rates");
      const PhysParams ph{c->phys.zeta, c->phys.v_turb, c->phys.omega, c->phys.grain_radius, c->phys.metallicity,
                          c->phys.g2d};
      if (launch_heating(c->net, ph, th->heat_index, nact, c->c_rho, c->c_T, c->c_Td, th->c_uv, c->c_ab, c->c_rate,
                         c->d_zeta ? c->c_zeta : nullptr, th->c_heat, s) != 0)
        return fail("heating");
      ThermalArgs ta{};
      ta.nact = nact;
      ta.idx = c->d_idx;
      ta.heat_c = th->c_heat;
      ta.heat = th->d_heat;
      ta.total_cool = th->d_total;
      ta.fmean = th->d_fmean;
      ta.fratio = th->d_fratio;
      ta.level_conv = th->level_conv;
      ta.first_time = th->first_time;
      ta.temp_fix = th->prm.temp_fix;
      ta.Tmin = th->prm.Tmin;
      ta.Tmax = th->prm.Tmax;
      ta.Fcrit = th->prm.Fcrit;
      ta.Tdiff = th->prm.Tdiff;
      ta.state = th->d_state;
      ta.T = c->d_T;
      ta.level_converged = th->d_lc;
      if (launch_thermal_update(ta, s) != 0) return fail("temperature update");
    }
  }
  // convergence (3DPDR.F90:1064-1099)
  if (th->level_conv && th->first_time) th->first_time = false;
  unsigned counts[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, (unsigned long long *)th->d_counts, 4);
  if (!ok(hipGetLastError(), "counts") || launch_count_converged(n, th->d_state, th->d_counts + 5, s) != 0 ||
      !ok(hipMemcpyAsync(&counts[5], th->d_counts + 5, 4, hipMemcpyDeviceToHost, s), "copy") ||
      !ok(hipStreamSynchronize(s), "sync"))
    return fail("convergence count");
  st.thermally_converged = (int)counts[5];
  if (th->level_conv) th->level_conv = false;
  if ((int)counts[5] == n) {
    th->done = true;
    st.done = 1;
  } else {
    // level populations: relative change of every coolant's populations
    // (3DPDR.F90:1086-1161); then the counts of label 1 and the update of the
    // populations (3DPDR.F90:1175-1228)
    if (launch_relch(t, n, c->d_ab, c->nspec, th->d_sol, th->d_pop, th->d_lc, th->d_cconv, th->d_counts, s) != 0 ||
        !ok(hipMemcpyAsync(counts, th->d_counts, 5 * 4, hipMemcpyDeviceToHost, s), "copy") ||
        !ok(hipStreamSynchronize(s), "sync"))
      return fail("level convergence");
    st.level_converged = (int)counts[0];
    for (int k = 0; k < 4; k++) st.coolant_converged[k] = (int)counts[1 + k];
    if ((int)counts[0] == n) {
      th->levpop_iteration = 0;
      th->level_conv = true;
    }
    for (int k = 0; k < 4; k++) th->pct[k] = (int)(100.0 * (double)counts[1 + k] / (double)n);
    const double levpop_pct = 100.0 * (double)counts[0] / (double)n;
    if ((int)levpop_pct >= 100)
      hipLaunchKernelGGL(fill_u8_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, th->d_lc, n);
    if (!ok(hipGetLastError(), "flags") ||
        launch_copy((long long)n * th->nlevtot, th->d_sol, th->d_pop, s) != 0 || !ok(hipStreamSynchronize(s), "sync"))
      return fail("population update");
  }
  st.levpop_iteration = th->levpop_iteration;
  st.level_conv = th->level_conv ? 1 : 0;
  st.first_time = th->first_time ? 1 : 0;
  st.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (out) *out = st;
  return 0;
}

const char *pdrchem_cloud_last_error(const pdrchem_cloud *c) {
  if (!c) return t_err.c_str();
  return c->err.c_str();
}

void pdrchem_cloud_destroy(pdrchem_cloud *c) { delete c; }

}  // extern "C"
