/*
 * Ray tracer of the column densities that feed the rate coefficients: CPU
 * restatement, TEST INFRASTRUCTURE ONLY (see pdr_oracle.h).
 *
 *   pdr_hp_pix2vec_nest   HEALPix nested pixel centre, healpix.F90:47-223
 *                         (with the zeroing of |v_x|,|v_y| < 1e-10, :172-173)
 *   pdr_hp_vec2pix_nest   vec2ang (healpix.F90:267-302) then ang2pix_nest_id
 *                         (:305-398): computed at ns_max = 8192 and degraded
 *   pdr_eval_points       evaluation points of every point of a point cloud
 *                         along the HEALPix rays, eval_points.F90:129-225
 *   pdr_ray_columns       Av per ray (CALC_UVFIELD, 3DPDR.F90:1650-1665) and
 *                         column densities of selected species
 *                         (calc_columndens, 3DPDR.F90:1818-1838)
 *   pdr_uvfield           attenuated field and the GUESS_TEMP temperature,
 *                         3DPDR.F90:1667-1708
 *
 * The candidate order of pdr_eval_points is either the reference's own
 * (PDR_ORDER_HEAPSORT: heapsort.F90 restated with its quirks -- the last
 * element is left out of the initial heap, and every value moved by a swap
 * passes through a single-precision temporary, as PR = SP without
 * -DDOUBLE_PRECISION, definitions.F90:14-18) or a strict order by (distance,
 * point index) (PDR_ORDER_STABLE), the order the device ray tracer uses.  The
 * two differ only where distances tie or round to the same float.
 *
 * Built with -ffp-contract=off (see Makefile): the operations and their order
 * follow the Fortran expressions, so results are bitwise those of the
 * reference on the same inputs.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pdr_oracle.h"

#define HP_HALFPI 1.570796326794896619231321691639751442099
#define HP_TWOPI 6.283185307179586476925286766559005768394
#define HP_TWOTHIRD 0.6666666666666666666666666666666666666666
#define HP_NS_MAX 8192
#define PDR_PC 3.08568025e18

/* bits b_p of v moved to positions 2p (x) -- mk_xy2pix, healpix.F90:3-44 */
static int spread_bits(int v) {
  int r = 0;
  for (int p = 0; v; p++, v >>= 1) r |= (v & 1) << (2 * p);
  return r;
}
/* inverse: even bits of k (x) or odd bits (y) -- mk_pix2xy, healpix.F90:225-264 */
static int gather_bits(int k, int odd) {
  int r = 0;
  k >>= odd;
  for (int p = 0; k; p++, k >>= 2) r |= (k & 1) << p;
  return r;
}

void pdr_hp_pix2vec_nest(int nside, int ipix, double v[3]) {
  static const int jrll[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
  static const int jpll[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};
  double fn = (double)nside;
  double fact1 = 1.0 / (3.0 * fn * fn);
  double fact2 = 2.0 / (3.0 * fn);
  int nl4 = 4 * nside;
  int npface = nside * nside;
  int face = ipix / npface;
  int ipf = ipix % npface;
  int ix = gather_bits(ipf, 0), iy = gather_bits(ipf, 1);
  int jrt = ix + iy, jpt = ix - iy;
  int jr = jrll[face] * nside - jrt - 1;
  int nr = nside;
  double z = (2 * nside - jr) * fact2;
  int kshift = ((jr - nside) % 2 + 2) % 2;
  if (jr < nside) {
    nr = jr;
    z = 1.0 - nr * nr * fact1;
    kshift = 0;
  } else if (jr > 3 * nside) {
    nr = nl4 - jr;
    z = -1.0 + nr * nr * fact1;
    kshift = 0;
  }
  int jp = (jpll[face] * nr + jpt + 1 + kshift) / 2;
  if (jp > nl4) jp -= nl4;
  if (jp < 1) jp += nl4;
  double phi = (jp - (kshift + 1) * 0.5) * (HP_HALFPI / nr);
  double sth = sqrt((1.0 - z) * (1.0 + z));
  v[0] = sth * cos(phi);
  v[1] = sth * sin(phi);
  v[2] = z;
  if (fabs(v[1]) < 1e-10) v[1] = 0.0;
  if (fabs(v[0]) < 1e-10) v[0] = 0.0;
}

int pdr_hp_vec2pix_nest(int nside, const double r[3]) {
  /* vec2ang */
  double dnorm = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  double theta = 0.0, phi = 0.0;
  if (dnorm != 0.0) {
    double zz = r[2] / dnorm;
    theta = acos(zz);
    if (r[0] != 0.0 || r[1] != 0.0) phi = atan2(r[1], r[0]);
    if (phi < 0.0) phi = phi + HP_TWOPI;
  }
  /* ang2pix_nest_id */
  double z = cos(theta);
  double za = fabs(z);
  double tt = (phi - floor(phi / HP_TWOPI) * HP_TWOPI) / HP_HALFPI; /* Fortran MODULO */
  int face, ix, iy;
  if (za <= HP_TWOTHIRD) {
    int jp = (int)(HP_NS_MAX * (0.5 + tt - z * 0.75));
    int jm = (int)(HP_NS_MAX * (0.5 + tt + z * 0.75));
    int ifp = jp / HP_NS_MAX, ifm = jm / HP_NS_MAX;
    if (ifp == ifm) face = (ifp % 4) + 4;
    else if (ifp < ifm) face = ifp % 4;
    else face = (ifm % 4) + 8;
    ix = jm % HP_NS_MAX;
    iy = HP_NS_MAX - (jp % HP_NS_MAX) - 1;
  } else {
    int ntt = (int)tt;
    if (ntt >= 4) ntt = 3;
    double tp = tt - ntt;
    double tmp = sqrt(3.0 * (1.0 - za));
    int jp = (int)(HP_NS_MAX * tp * tmp);
    int jm = (int)(HP_NS_MAX * (1.0 - tp) * tmp);
    if (jp > HP_NS_MAX - 1) jp = HP_NS_MAX - 1;
    if (jm > HP_NS_MAX - 1) jm = HP_NS_MAX - 1;
    if (z >= 0) {
      face = ntt;
      ix = HP_NS_MAX - jm - 1;
      iy = HP_NS_MAX - jp - 1;
    } else {
      face = ntt + 8;
      ix = jp;
      iy = jm;
    }
  }
  int ipf = (spread_bits(ix / 128) + 2 * spread_bits(iy / 128)) * (128 * 128) +
            (spread_bits(ix % 128) + 2 * spread_bits(iy % 128));
  int d = HP_NS_MAX / nside;
  ipf = ipf / (d * d);
  return ipf + face * nside * nside;
}

/* heapsort.F90:10-90 on 1-based plist[1..n], rlist[1..n], as the reference
 * runs it: the initial heap covers 1..n-1 only, and a value moved through the
 * swap temporary is rounded to single precision. */
static void ref_heapsort(int n, int *plist, double *rlist) {
  for (int start = n / 2; start > 0; start--) {
    int sift_end = n - 1, root = start;
    for (;;) {
      int child = root * 2;
      if (child > sift_end) break;
      if (child < sift_end && rlist[child] < rlist[child + 1]) child++;
      if (rlist[root] < rlist[child]) {
        float sr = (float)rlist[root];
        rlist[root] = rlist[child];
        rlist[child] = (double)sr;
        int si = plist[root]; plist[root] = plist[child]; plist[child] = si;
        root = child;
      } else {
        break;
      }
    }
  }
  for (int end = n; end > 1;) {
    float sr = (float)rlist[end];
    rlist[end] = rlist[1];
    rlist[1] = (double)sr;
    int si = plist[end]; plist[end] = plist[1]; plist[1] = si;
    end--;
    int sift_end = end, root = 1;
    for (;;) {
      int child = root * 2;
      if (child > sift_end) break;
      if (child < sift_end && rlist[child] < rlist[child + 1]) child++;
      if (rlist[root] < rlist[child]) {
        float s2 = (float)rlist[root];
        rlist[root] = rlist[child];
        rlist[child] = (double)s2;
        int sj = plist[root]; plist[root] = plist[child]; plist[child] = sj;
        root = child;
      } else {
        break;
      }
    }
  }
}

typedef struct { double r; int id; } dist_id;
static int cmp_dist_id(const void *a, const void *b) {
  const dist_id *x = (const dist_id *)a, *y = (const dist_id *)b;
  if (x->r < y->r) return -1;
  if (x->r > y->r) return 1;
  return (x->id > y->id) - (x->id < y->id);
}

/*
 * Evaluation points of every point of a cloud (eval_points.F90:129-225; the
 * dark element branch :15-110 is the same loop for one more origin).
 * xyz[npts][3] positions (pc), etype[npts] (2 = ionised: a ray stops at it).
 * Outputs, with M = maxpoints + 1 slots per ray:
 *   epray[npts][nrays]           evaluation points per ray
 *   proj[npts][nrays][M]         projected point of evaluation point i
 *                                (0-based ids; slot 0 = the origin itself)
 *   epoint[npts][nrays][M][3]    evaluation point coordinates (slot 0 = origin)
 * Returns 0, or -1 when a ray needs more than maxpoints points (the
 * reference STOPs there, eval_points.F90:210).
 */
int pdr_eval_points(int npts, const double *xyz, const int *etype, int level, double theta_crit, int order,
                    int maxpoints, int *epray, int *proj, double *epoint) {
  int nside = 1 << level, nrays = 12 * nside * nside, M = maxpoints + 1;
  double *vec = (double *)malloc(sizeof(double) * 3 * nrays);
  for (int j = 0; j < nrays; j++) pdr_hp_pix2vec_nest(nside, j, vec + 3 * j);
  int n1 = npts - 1;
  int *rb = (int *)malloc(sizeof(int) * (npts + 1));
  double *ra = (double *)malloc(sizeof(double) * (npts + 1));
  dist_id *di = (dist_id *)malloc(sizeof(dist_id) * (npts + 1));
  double *ep = (double *)malloc(sizeof(double) * 3 * nrays);
  char *kill = (char *)malloc(nrays);
  int status = 0;
  for (int p = 0; p < npts && status == 0; p++) {
    const double *o = xyz + 3 * p;
    memset(kill, 0, nrays);
    int kk = 0;
    for (int i = 0; i < npts; i++) {
      if (i == p) continue;
      kk++;
      double rv[3] = {xyz[3 * i] - o[0], xyz[3 * i + 1] - o[1], xyz[3 * i + 2] - o[2]};
      ra[kk] = sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
      rb[kk] = i;
    }
    if (order == PDR_ORDER_HEAPSORT) {
      ref_heapsort(n1, rb, ra);
    } else {
      for (int k = 1; k <= n1; k++) { di[k - 1].r = ra[k]; di[k - 1].id = rb[k]; }
      qsort(di, n1, sizeof(dist_id), cmp_dist_id);
      for (int k = 1; k <= n1; k++) { ra[k] = di[k - 1].r; rb[k] = di[k - 1].id; }
    }
    double radius = ra[n1];
    memset(ep, 0, sizeof(double) * 3 * nrays);
    for (int j = 0; j < nrays; j++) {
      epray[p * nrays + j] = 0;
      proj[(size_t)(p * nrays + j) * M] = p;
      double *e0 = epoint + ((size_t)(p * nrays + j) * M) * 3;
      e0[0] = o[0]; e0[1] = o[1]; e0[2] = o[2];
    }
    for (int k = 1; k <= n1; k++) {
      int id = rb[k];
      double rv[3] = {xyz[3 * id] - o[0], xyz[3 * id + 1] - o[1], xyz[3 * id + 2] - o[2]};
      int ipix = pdr_hp_vec2pix_nest(nside, rv);
      if (kill[ipix]) continue;
      double h[3], *e = ep + 3 * ipix;
      for (int c = 0; c < 3; c++) h[c] = 1.1 * radius * vec[3 * ipix + c];
      double a[3] = {rv[0] - e[0], rv[1] - e[1], rv[2] - e[2]};
      double b[3] = {h[0] - e[0], h[1] - e[1], h[2] - e[2]};
      double dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
      double na = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
      double nb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
      double ang = acos(dot / (na * nb));
      if (!(ang <= theta_crit)) continue;
      if (h[2] != 0.0) {
        e[2] = (h[0] * h[2] * rv[0] + h[1] * h[2] * rv[1] + (h[2] * h[2]) * rv[2]) /
               (h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
        e[1] = e[2] * h[1] / h[2];
        e[0] = e[2] * h[0] / h[2];
      } else if (h[0] == 0.0) {
        e[0] = 0.0; e[1] = rv[1]; e[2] = 0.0;
      } else if (h[1] == 0.0) {
        e[0] = rv[0]; e[1] = 0.0; e[2] = 0.0;
      } else {
        e[2] = 0.0;
        e[0] = ((h[0] * h[0]) * rv[0] + h[0] * h[1] * rv[1]) / (h[0] * h[0] + h[1] * h[1]);
        e[1] = e[0] * h[1] / h[0];
      }
      int n = ++epray[p * nrays + ipix];
      if (n > maxpoints) { status = -1; break; }
      size_t s = (size_t)(p * nrays + ipix) * M + n;
      proj[s] = id;
      epoint[3 * s] = e[0] + o[0];
      epoint[3 * s + 1] = e[1] + o[1];
      epoint[3 * s + 2] = e[2] + o[2];
      if (etype && etype[id] == 2) kill[ipix] = 1;
    }
  }
  free(vec); free(rb); free(ra); free(di); free(ep); free(kill);
  return status;
}

/*
 * Column densities along every ray of every point, from the evaluation
 * points: Av (CALC_UVFIELD, 3DPDR.F90:1650-1664: trapezoid of n_H times the
 * step, times pc and AV_fac) and the columns of nsel selected species
 * (calc_columndens, 3DPDR.F90:1818-1838: trapezoid of n_H x_k).
 * rho[npts]; xsel[npts][nsel] abundances of the selected species (NULL when
 * nsel = 0).  Out: Av[npts][nrays], cols[npts][nrays][nsel].
 */
void pdr_ray_columns(int npts, int nrays, int maxpoints, const int *epray, const int *proj,
                     const double *epoint, const double *rho, int nsel, const double *xsel, double av_fac,
                     double *Av, double *cols) {
  int M = maxpoints + 1;
  for (int p = 0; p < npts; p++)
    for (int j = 0; j < nrays; j++) {
      size_t b = (size_t)(p * nrays + j) * M;
      double cd = 0.0;
      double *cj = cols ? cols + (size_t)(p * nrays + j) * nsel : 0;
      for (int k = 0; k < nsel; k++) cj[k] = 0.0;
      for (int i = 1; i <= epray[p * nrays + j]; i++) {
        const double *e0 = epoint + 3 * (b + i - 1), *e1 = epoint + 3 * (b + i);
        double step = sqrt((e0[0] - e1[0]) * (e0[0] - e1[0]) + (e0[1] - e1[1]) * (e0[1] - e1[1]) +
                           (e0[2] - e1[2]) * (e0[2] - e1[2]));
        int q0 = proj[b + i - 1], q1 = proj[b + i];
        cd = cd + ((rho[q0] + rho[q1]) / 2.) * step * PDR_PC;
        for (int k = 0; k < nsel; k++)
          cj[k] = cj[k] + step * PDR_PC * (rho[q0] * xsel[q0 * nsel + k] + rho[q1] * xsel[q1 * nsel + k]) / 2.;
      }
      Av[p * nrays + j] = cd * av_fac;
    }
}

/*
 * Per-ray surface field and attenuated field of every point
 * (CALC_UVFIELD, 3DPDR.F90:1667-1694) and the GUESS_TEMP gas temperature
 * (:1698-1700).  fieldchoice 0 = "ISO" (gext[0]/nrays on every ray), 1 =
 * "UNI" (only the ray most opposed to gext receives -gext.v).  Out:
 * G[npts][nrays], uv[npts], tguess[npts] (either may be NULL).
 */
void pdr_uvfield(int npts, int level, int fieldchoice, const double gext[3], const double *Av, double uv_fac,
                 double *G, double *uv, double *tguess) {
  int nside = 1 << level, nrays = 12 * nside * nside;
  double *vec = (double *)malloc(sizeof(double) * 3 * nrays);
  for (int j = 0; j < nrays; j++) pdr_hp_pix2vec_nest(nside, j, vec + 3 * j);
  int jmin = 0;
  double best = 0.0;
  for (int j = 0; j < nrays; j++) {
    const double *v = vec + 3 * j;
    double d = -(gext[0] * v[0] + gext[1] * v[1] + gext[2] * v[2]);
    if (d >= best) { best = d; jmin = j; }
  }
  for (int p = 0; p < npts; p++) {
    double u = 0.0;
    for (int j = 0; j < nrays; j++) {
      const double *v = vec + 3 * j;
      double g = 0.0;
      if (fieldchoice == 0) g = gext[0] / (double)nrays;
      else if (j == jmin) g = -(gext[0] * v[0] + gext[1] * v[1] + gext[2] * v[2]);
      G[p * nrays + j] = g;
      u = u + g * exp(-Av[p * nrays + j] * uv_fac);
      if (u < 1.0e-50) u = 0.0;
    }
    if (uv) uv[p] = u;
    if (tguess) tguess[p] = 10.0 * (1.0 + pow(2. * u, 1.0 / 3.0));
  }
  free(vec);
}
