// Batched variable-order BDF, two stiff systems per wavefront (wave64): one
// per 32-lane half, lane = row of its system, for networks of up to 32
// equations (the reduced network).
//
// Included once per network inside a namespace that defines `namespace NET`
// (a generated csrc/gen/net_<name>.h, whose P_* tables are this layout's).
// The integrator is bdf_core.h, as in the wavefront-per-system kernel
// (bdf_wave_impl.h), with the half-wave as its team.  In that kernel a
// system's step control -- the coefficients, norms, step and order choice,
// the Newton bookkeeping -- is uniform scalar work that occupies a whole
// wavefront's instruction for one system; here the same instruction carries
// both halves' systems, and the linear algebra packs a full row per lane
// (32 columns) instead of half a row.
//
// The arithmetic is the wavefront kernel's, operation for operation: the
// same flux products, the same RHS chunks summed in the same order (the
// generator packs the wavefront layout's chunks into this layout's cells),
// the same Jacobian term order, pivot choice, Gauss-Jordan updates and
// inverse-apply association, and the same integrator code, so both kernels
// return the same abundances bit for bit.
//
// Each half runs its own sequence of points: the integrator's output loop
// (calculate_abundances.c:214-305) is flattened into one loop of steps in
// which a half that has finished a CVode call, or a whole point, moves on
// (next output time, next point from the queue) while the other half keeps
// stepping, so the halves stay in step with each other except inside a step
// (Newton iterations, setups, error-test retries) and at the very end.

#include "wave_ops.h"

using namespace pdrw;

// Within a half every lane holds the same integrator scalars, but the two
// halves hold different systems': the integrator's reads of its shared
// record stay per lane (the wavefront kernel makes them wave-uniform).
__device__ __forceinline__ int uni(int v) { return v; }
__device__ __forceinline__ double uni(double v) { return v; }

#include "bdf_core.h"

static_assert(NET::NEQ <= 32, "one row per lane of a half");

#ifndef PDR_PAIR_WPE32
#define PDR_PAIR_WPE32 2
#endif
#ifndef PDR_PAIR_WPE64
#define PDR_PAIR_WPE64 2
#endif
template <typename MT>
constexpr int kPairWavesPerEU = sizeof(MT) == 4 ? PDR_PAIR_WPE32 : PDR_PAIR_WPE64;

// slot rate constants in LDS for the FP32 form, in the workspace (L2) for
// the FP64 form (whose flux partials take twice the LDS)
#ifndef PDR_PAIR_KP_LDS32
#define PDR_PAIR_KP_LDS32 1
#endif
template <typename MT>
constexpr bool kPKpInLDS = sizeof(MT) == 4 && PDR_PAIR_KP_LDS32;

// per-system workspace in device memory: the saved Jacobian ([column][lane])
// and, FP64 form, the slot rate constants
template <typename MT>
struct PairWS {
  MT J[32 * 32];
  double Kp[kPKpInLDS<MT> ? 1 : NET::P_NSLOT];
};

template <typename MT>
constexpr int kPFWords = (NET::P_NSLOT + 1) > (int)(((2 * NET::P_NSLOT + 1) * sizeof(MT) + 7) / 8)
                             ? (NET::P_NSLOT + 1)
                             : (int)(((2 * NET::P_NSLOT + 1) * sizeof(MT) + 7) / 8);

// per-system shared memory (one record per half)
template <typename MT>
struct SysLDS {
  double X[64];                    // factor vector: y, x_e, 1
  double F[kPFWords<MT>];          // fluxes (RHS) or flux partials (Jacobian)
  double Kp[kPKpInLDS<MT> ? NET::P_NSLOT : 1];
  double Zh[(QMAX - 1) * 32];      // Nordsieck entries zn[2..QMAX] of each lane
  union {
    struct {
      double V[32];                // vector exchange (M build, solve)
      double V2[32];
    };
    double P[NET::P_NSEG * 32 + 1];  // RHS cell partials; P[P_NSEG*32] = 0
    alignas(16) MT PR[32];         // pivot row of the Gauss-Jordan step
  };
  BdfU u;
};

template <typename MT, int NW>
struct PairBlockLDS {
  uint32_t rterm[NET::P_TM * 32];
  uint32_t fab[NET::P_NSLOT];
  SysLDS<MT> s[2 * NW];
};

// per-lane integrator vectors and matrix row
template <typename MT>
struct Bdf2 {
  double zn[QMAX + 1];
  double ewt, acor, y, ftemp, wsc;
  MT A[32];
  int invp;
  double LDSA *zh;  // zh[(j - 2) * 32] = zn[j] of this lane
  __device__ __forceinline__ void hist_in() {
#pragma unroll
    for (int j = 2; j <= QMAX; j++) zn[j] = zh[(j - 2) * 32];
  }
  __device__ __forceinline__ void hist_out() {
#pragma unroll
    for (int j = 2; j <= QMAX; j++) zh[(j - 2) * 32] = zn[j];
  }
};

// sum / max over each 32-lane half, the half's result in each of its lanes;
// the same association as the wavefront kernel's wave_sum_lo
__device__ __forceinline__ double half_sum(double v, int hh) {
  v = row_sum(v);
  const double s0 = rdlane(v, 0) + rdlane(v, 16);
  const double s1 = rdlane(v, 32) + rdlane(v, 48);
  return hh ? s1 : s0;
}
__device__ __forceinline__ double half_max_d(double v, int hh) {
  v = fmax(v, dpp<DPP_XOR1>(v));
  v = fmax(v, dpp<DPP_XOR2>(v));
  v = fmax(v, dpp<DPP_HALF_MIRROR>(v));
  v = fmax(v, dpp<DPP_MIRROR>(v));
  const double m0 = fmax(rdlane(v, 0), rdlane(v, 16));
  const double m1 = fmax(rdlane(v, 32), rdlane(v, 48));
  return hh ? m1 : m0;
}
__device__ __forceinline__ uint32_t half_max_key(uint32_t k, int hh) {
  k = row_max(k);
  const uint32_t m0 = max(rdlane_u(k, 0), rdlane_u(k, 16));
  const uint32_t m1 = max(rdlane_u(k, 32), rdlane_u(k, 48));
  return hh ? m1 : m0;
}

// the half-wavefront as the integrator's team (see bdf_core.h)
template <typename MT>
struct Ctx2 {
  int L, i, hh;
  bool row_ok;
  double q;                    // charge of row i
  uint32_t cells[2];           // the row's partial indices, four bytes each
  const uint32_t LDSA *rterm;  // block LDS copies of the RHS tables
  const uint32_t LDSA *fab;
  SysLDS<MT> LDSA *S;
  PairWS<MT> *W;
  double rtol, atol;

  __device__ __forceinline__ bool lead() const { return (L & 31) == 0; }
  __device__ __forceinline__ bool own() const { return row_ok; }
  __device__ __forceinline__ double sum(double v) { return half_sum(v, hh); }
  __device__ __forceinline__ double sum_own_sq(double v) { return half_sum(v, hh); }
  __device__ __forceinline__ double vmax(double v) { return half_max_d(v, hh); }
  __device__ __forceinline__ void sync() { wsync(); }
  __device__ __forceinline__ double rhs(double y, double &xe);
  __device__ __forceinline__ int setup(BdfU LDSA &u, Bdf2<MT> &b, int convfail);
  __device__ __forceinline__ double solve(const BdfU LDSA &u, const Bdf2<MT> &b, double r);
};

template <typename MT>
__device__ __forceinline__ double kp2(const Ctx2<MT> &c, int s) {
  if constexpr (kPKpInLDS<MT>) return c.S->Kp[s];
  else return c.W->Kp[s];
}

// ---------------------------------------------------------------------
// RHS: ydot_i of this lane's row; x_e of the half via xe_out
// ---------------------------------------------------------------------
template <typename MT>
__device__ __forceinline__ double pair_rhs(Ctx2<MT> &c, double y, double &xe_out) {
  SysLDS<MT> LDSA &S = *c.S;
  if (c.row_ok) S.X[c.i] = y;
  const double xe = half_sum(c.row_ok ? c.q * y : 0.0, c.hh);
  xe_out = xe;
  if (c.lead()) S.X[NET::NEQ] = xe;
  wsync();
  uint32_t fv[NET::P_ROUNDS];
  double xa[NET::P_ROUNDS], xb[NET::P_ROUNDS];
#pragma unroll
  for (int r = 0; r < NET::P_ROUNDS; r++) fv[r] = c.fab[r * 32 + c.i];
#pragma unroll
  for (int r = 0; r < NET::P_ROUNDS; r++) {
    xa[r] = S.X[fv[r] & 0xff];
    xb[r] = S.X[(fv[r] >> 8) & 0xff];
  }
#pragma unroll
  for (int r = 0; r < NET::P_ROUNDS; r++) S.F[r * 32 + c.i] = kp2(c, r * 32 + c.i) * xa[r] * xb[r];
  wsync();
  // each segment: its term words, then their gathers, then the products
  // (in groups of G, as the wavefront kernel), into the cell's partial
  constexpr int G = 8;
  const unsigned char LDSA *Fb = (const unsigned char LDSA *)S.F;
  double part[NET::P_NSEG];
  int t0 = 0;
#pragma unroll
  for (int k = 0; k < NET::P_NSEG; k++) {
    double acc = 0.0;
#pragma unroll
    for (int g = 0; g < NET::P_SEG[k]; g += G) {
      uint32_t w[G];
      double f[G];
#pragma unroll
      for (int u = 0; u < G; u++)
        if (g + u < NET::P_SEG[k]) w[u] = c.rterm[(t0 + g + u) * 32 + c.i];
#pragma unroll
      for (int u = 0; u < G; u++)
        if (g + u < NET::P_SEG[k]) f[u] = *(const double LDSA *)(Fb + (w[u] & 0xffff));
#pragma unroll
      for (int u = 0; u < G; u++)
        if (g + u < NET::P_SEG[k]) acc = fma(__hiloint2double((int)(w[u] & 0xffff0000u), 0), f[u], acc);
    }
    part[k] = acc;
    t0 += NET::P_SEG[k];
  }
#pragma unroll
  for (int k = 0; k < NET::P_NSEG; k++) S.P[k * 32 + c.i] = part[k];
  wsync();
  double yd = 0.0;
#pragma unroll
  for (int u = 0; u < NET::P_PM; u++) yd += S.P[(c.cells[u >> 2] >> (8 * (u & 3))) & 0xff];
  wsync();
  return yd;
}

// ---------------------------------------------------------------------
// exact Jacobian at the factor vector in S.X: flux partials, then the two
// passes of the per-column term lists (columns 0-15 and x_e, then 16-31);
// row i of J = acc + g_i * charge, saved to the workspace
// ---------------------------------------------------------------------
template <typename MT>
__device__ __forceinline__ void pair_jacobian(Ctx2<MT> &c, MT (&jv)[32]) {
  SysLDS<MT> LDSA &S = *c.S;
  MT LDSA *D = (MT LDSA *)S.F;
#pragma unroll
  for (int r = 0; r < NET::P_ROUNDS; r++) {
    const int s = r * 32 + c.i;
    const uint32_t f = c.fab[s];
    const double kp = kp2(c, s);
    D[2 * s] = (MT)(kp * S.X[(f >> 8) & 0xff]);
    D[2 * s + 1] = (MT)(kp * S.X[f & 0xff]);
  }
  wsync();
  constexpr int G = 8;
  const uint32_t *jt = NET::p_jac_term + c.i;
  double g_e;
  {
    double acc[17];
#pragma unroll
    for (int j = 0; j < 17; j++) acc[j] = 0.0;
#pragma unroll
    for (int g = 0; g < NET::P_JA; g += G) {
      uint32_t w[G];
      MT dv[G];
#pragma unroll
      for (int u = 0; u < G; u++)
        if (g + u < NET::P_JA) w[u] = jt[(g + u) * 32];
#pragma unroll
      for (int u = 0; u < G; u++)
        if (g + u < NET::P_JA) dv[u] = D[w[u] & 0xffff];
#pragma unroll
      for (int u = 0; u < G; u++)
        if (g + u < NET::P_JA) {
          const int j = NET::p_jac_acc[g + u];
          acc[j] = fma((double)(int16_t)(w[u] >> 16), (double)dv[u], acc[j]);
        }
    }
    g_e = acc[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
      jv[j] = (MT)fma(g_e, (double)NET::charge[j], acc[j]);
      c.W->J[j * 32 + c.i] = jv[j];
    }
  }
  {
    double acc[16];
#pragma unroll
    for (int j = 0; j < 16; j++) acc[j] = 0.0;
#pragma unroll
    for (int g = NET::P_JA; g < NET::P_JT; g += G) {
      uint32_t w[G];
      MT dv[G];
#pragma unroll
      for (int u = 0; u < G; u++)
        if (g + u < NET::P_JT) w[u] = jt[(g + u) * 32];
#pragma unroll
      for (int u = 0; u < G; u++)
        if (g + u < NET::P_JT) dv[u] = D[w[u] & 0xffff];
#pragma unroll
      for (int u = 0; u < G; u++)
        if (g + u < NET::P_JT) {
          const int j = NET::p_jac_acc[g + u];
          acc[j] = fma((double)(int16_t)(w[u] >> 16), (double)dv[u], acc[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < 16; j++) {
      jv[16 + j] = (MT)fma(g_e, (double)NET::charge[16 + j], acc[j]);
      c.W->J[(16 + j) * 32 + c.i] = jv[16 + j];
    }
  }
  wsync();
  // in the FP64 form the partials covered the RHS's zero flux F[P_NSLOT]
  if (c.lead()) S.F[NET::P_NSLOT] = 0.0;
}

// ---------------------------------------------------------------------
// Gauss-Jordan inverse (in place, partial pivoting) of the scaled matrix,
// the wavefront kernel's steps with a whole row per lane: the pivot is the
// unused row of the half with the largest |a[i][k]| (packed keys, one DPP
// max), its row goes through the half's LDS record, and every lane updates
// its 32 entries.  Rows are not moved: A holds the inverse in pivot order
// and invp is the step at which this row was the pivot.  False if singular.
// ---------------------------------------------------------------------
template <typename MT>
__device__ __forceinline__ bool pair_gj(const Ctx2<MT> &c, MT (&A)[32], int &invp) {
  SysLDS<MT> LDSA &S = *c.S;
  bool used = !c.row_ok;
  invp = c.i;
#pragma unroll
  for (int k = 0; k < NET::NEQ; k++) {
    const MT colk = A[k];
    const uint32_t key = !used ? pivot_key<5>(fabs(colk), c.i) : 0u;
    bool ok;
    const int p = pivot_row<5, MT>(half_max_key(key, c.hh), ok);
    if (!ok) return false;
    const bool isp = (c.i == p);
    if (isp) {
#pragma unroll
      for (int j = 0; j < 32; j++) S.PR[j] = A[j];
    }
    wsync();
    MT pr[32];
#pragma unroll
    for (int j = 0; j < 32; j++) pr[j] = S.PR[j];
    wsync();
    const MT piv = pr[k];
    MT d;
    if constexpr (sizeof(MT) == 4) d = __builtin_amdgcn_rcpf(piv);
    else d = (MT)1.0 / piv;
    used = used || isp;
    invp = isp ? k : invp;
    const MT g = isp ? d : -colk * d;
#pragma unroll
    for (int j = 0; j < 32; j++) A[j] = fma(g, pr[j], isp ? (MT)0 : A[j]);
    A[k] = g;
  }
  return true;
}

// x = M^-1 b with the pivot-ordered inverse: columns 0-15 and 16-31 summed
// apart and then added, as the wavefront kernel's two halves
template <typename MT>
__device__ __forceinline__ double pair_inv_apply(const Ctx2<MT> &c, const MT (&A)[32], int invp, double b) {
  SysLDS<MT> LDSA &S = *c.S;
  S.V[invp] = c.row_ok ? b : 0.0;
  wsync();
  double lo = 0.0, hi = 0.0;
#pragma unroll
  for (int j = 0; j < 16; j++) lo = fma((double)A[j], S.V[j], lo);
#pragma unroll
  for (int j = 0; j < 16; j++) hi = fma((double)A[16 + j], S.V[16 + j], hi);
  S.V2[invp] = lo + hi;
  wsync();
  const double x = S.V2[c.i];
  wsync();
  return c.row_ok ? x : 0.0;
}

template <typename MT>
__device__ __forceinline__ int pair_lsetup(Ctx2<MT> &c, BdfU LDSA &u, Bdf2<MT> &b, int convfail) {
  SysLDS<MT> LDSA &S = *c.S;
#pragma unroll
  for (int j = 0; j < 32; j++) b.A[j] = c.W->J[j * 32 + c.i];
  const double dgamma = fabs(u.gamma / u.gammap - 1.0);
  const bool jbad = (u.nst == 0) || (u.nst > u.nstlj + MSBJ) ||
                    ((convfail == CV_FAIL_BAD_J) && !(dgamma >= DGMAX_J)) || (convfail == CV_FAIL_OTHER);
  if (jbad) {
    const int nst = u.nst;
    if (c.lead()) {
      u.nje++;
      u.nstlj = nst;
      u.jcur = 1;
    }
    if (c.row_ok) S.X[c.i] = b.zn[0];
    const double xe = half_sum(c.row_ok ? c.q * b.zn[0] : 0.0, c.hh);
    if (c.lead()) S.X[NET::NEQ] = xe;
    wsync();
    pair_jacobian(c, b.A);
  } else {
    if (c.lead()) u.jcur = 0;
  }
  const double gamma = u.gamma;
  b.wsc = c.row_ok ? b.ewt : 1.0;
  S.V[c.i] = 1.0 / b.wsc;
  wsync();
#pragma unroll
  for (int j = 0; j < 32; j++) {
    const double wjinv = S.V[j];
    double m = -gamma * (double)b.A[j];
    if (j == c.i) m += 1.0;
    if (!c.row_ok) m = (j == c.i) ? 1.0 : 0.0;
    b.A[j] = (MT)(m * b.wsc * wjinv);
  }
  wsync();
  const bool ok = pair_gj(c, b.A, b.invp);
  return ok ? 0 : 1;
}

template <typename MT>
__device__ __forceinline__ double pair_lsolve(Ctx2<MT> &c, const BdfU LDSA &u, const Bdf2<MT> &b, double r) {
  const double gamrat = u.gamrat;
  double x = pair_inv_apply(c, b.A, b.invp, r * b.wsc) / b.wsc;
  if (gamrat != 1.0) x *= 2.0 / (1.0 + gamrat);
  return c.row_ok ? x : 0.0;
}

template <typename MT>
__device__ __forceinline__ double Ctx2<MT>::rhs(double y, double &xe) { return pair_rhs(*this, y, xe); }
template <typename MT>
__device__ __forceinline__ int Ctx2<MT>::setup(BdfU LDSA &u, Bdf2<MT> &b, int convfail) {
  return pair_lsetup(*this, u, b, convfail);
}
template <typename MT>
__device__ __forceinline__ double Ctx2<MT>::solve(const BdfU LDSA &u, const Bdf2<MT> &b, double r) {
  return pair_lsolve(*this, u, b, r);
}

// rates of one point -> the half's flux constants
template <typename MT>
__device__ __forceinline__ void pair_load_kp(Ctx2<MT> &c, const double *rate, double nH) {
#pragma unroll
  for (int r = 0; r < NET::P_ROUNDS; r++) {
    const int s = r * 32 + c.i;
    const int re = NET::p_slot_reac[s];
    const uint32_t f = c.fab[s];
    const double k = (re >= 0) ? rate[re] : 0.0;
    const double kps = ((f >> 16) & 1) ? k * nH : k;
    if constexpr (kPKpInLDS<MT>) c.S->Kp[s] = kps;
    else c.W->Kp[s] = kps;
  }
}

// phases of a half's run through its points
enum { PH_FETCH = 0, PH_CALL = 1, PH_STEP = 2, PH_END = 3, PH_EXIT = 4 };

// the half's share of a uniform value taken by its lead lane
__device__ __forceinline__ int half_bcast(int v, int hh) {
  const int v0 = __builtin_amdgcn_readlane(v, 0);
  const int v1 = __builtin_amdgcn_readlane(v, 32);
  return hh ? v1 : v0;
}

// persistent kernel: each half of each wavefront pulls grid points from the
// queue and integrates them one after another (calculate_abundances.c:
// 132-326 per point, output loop :214-305 flattened into the step loop)
template <typename MT, int NW>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(kPairWavesPerEU<MT>)))
bdf_pair_kernel(PointArgs a, PairWS<MT> *ws) {
  __shared__ PairBlockLDS<MT, NW> blk;
  for (int k = threadIdx.x; k < NET::P_TM * 32; k += blockDim.x) blk.rterm[k] = NET::p_rhs_term[k];
  for (int k = threadIdx.x; k < NET::P_NSLOT; k += blockDim.x) blk.fab[k] = NET::p_slot_fab[k];
  __syncthreads();
  Ctx2<MT> c;
  c.L = lane_id();
  c.i = c.L & 31;
  c.hh = c.L >> 5;
  c.row_ok = c.i < NET::NEQ;
  c.q = (double)NET::charge[c.i];
  c.cells[0] = c.cells[1] = 0u;
#pragma unroll
  for (int u = 0; u < NET::P_PM; u++) c.cells[u >> 2] |= (uint32_t)NET::p_row_cells[c.i * NET::P_PM + u] << (8 * (u & 3));
  c.rterm = (const uint32_t LDSA *)blk.rterm;
  c.fab = (const uint32_t LDSA *)blk.fab;
  const int wave = threadIdx.x >> 6;
  c.S = (SysLDS<MT> LDSA *)&blk.s[2 * wave + c.hh];
  c.W = ws + ((size_t)blockIdx.x * NW + wave) * 2 + c.hh;
  c.rtol = a.rtol;
  c.atol = a.atol;
  BdfU LDSA &u = c.S->u;
  SysLDS<MT> LDSA &S = *c.S;
  S.P[NET::P_NSEG * 32] = 0.0;  // the zero partial (no other phase writes it)
  Bdf2<MT> b;
  b.zh = S.Zh + c.i;
  int phase = PH_FETCH, sys = 0, nfails = 0, status = 0, flag = 0;
  long long nstloc = 0;
  bool first = true;
  double tout = 0.0, t = 0.0, tend = 0.0, yout = 0.0;
  double *ab = nullptr;
  for (;;) {
    if (phase == PH_FETCH) {
      int got = 0;
      if (c.lead()) got = (int)atomicAdd(a.counter, 1u);
      got = half_bcast(got, c.hh);
      if ((unsigned int)got >= (unsigned int)a.nsys) {
        phase = PH_EXIT;
      } else {
        sys = got;
        pair_load_kp(c, a.rate + (size_t)sys * a.rate_stride, a.density[sys]);
        ab = a.abundance + (size_t)sys * a.ab_stride;
        const double y0 = c.row_ok ? ab[c.i] : 0.0;
        const double xe0 = ab[a.nelect0];
        wsync();
        S.X[c.i] = (c.i == NET::NEQ + 1) ? 1.0 : 0.0;
        S.X[32 + c.i] = (32 + c.i == NET::NEQ + 1) ? 1.0 : 0.0;
        if (c.lead()) {
          S.F[NET::P_NSLOT] = 0.0;
          ((MT LDSA *)S.F)[2 * NET::P_NSLOT] = (MT)0;
          bdf_init_uniform(u, a.t0, xe0);
        }
        wsync();
        b.zn[0] = y0;
#pragma unroll
        for (int j = 1; j <= QMAX; j++) b.zn[j] = 0.0;
        b.hist_out();
        b.acor = 0.0; b.y = y0; b.ftemp = 0.0; b.wsc = 1.0; b.invp = c.i;
#pragma unroll
        for (int j = 0; j < 32; j++) b.A[j] = (MT)0;
        b.ewt = 1.0 / (c.rtol * fabs(y0) + c.atol);
        tend = a.tend_pts ? a.tend_pts[sys] : a.tend;
        tout = a.tout0;
        t = a.t0;
        yout = y0;
        nfails = 0;
        status = 0;
        flag = 0;
        first = true;
        phase = PH_CALL;
      }
    }
    // the wavefront leaves once both halves have run out of points (every
    // lane is active here: the phases' branches have rejoined)
    if (__ballot(phase != PH_EXIT) == 0ull) break;
    if (phase == PH_CALL) {
      // ---- start of CVode(tout, CV_NORMAL)
      flag = 0;
      if (first) {
        b.zn[1] = f_eval(c, u, b.zn[0]);
        if (c.lead()) u.nfe++;
        c.sync();
        const int hf = hin(c, u, b, tout);
        if (hf != 0) {
          flag = hf;
          t = u.tn;
          yout = b.zn[0];
          phase = PH_END;
        } else {
          const double h = u.h;
          c.sync();
          if (c.lead()) { u.hscale = h; u.hprime = h; }
          c.sync();
          b.zn[1] *= h;
          first = false;
          nstloc = 0;
          phase = PH_STEP;
        }
      } else if ((u.tn - tout) * u.h >= 0.0) {
        t = tout;
        yout = get_dky(u, b, tout);
        phase = PH_END;
      } else {
        nstloc = 0;
        phase = PH_STEP;
      }
    }
    if (phase == PH_STEP) {
      // ---- one step of the CVode call
      if (u.nst > 0) b.ewt = 1.0 / (c.rtol * fabs(b.zn[0]) + c.atol);
      int why = 0;  // 1 = tout reached, < 0 = failure flag
      if (nstloc >= a.mxstep) {
        why = -1;
      } else {
        const double nrm = wrms(c, b.zn[0], b.ewt);
        if (UROUND * nrm > 1.0) {
          why = -2;
        } else {
          const int kflag = bdf_step(c, u, b);
          if (kflag != 0) {
            why = kflag;
          } else {
            nstloc++;
            if ((u.tn - tout) * u.h >= 0.0) why = 1;
          }
        }
      }
      if (why == 1) {
        t = tout;
        yout = get_dky(u, b, tout);
        phase = PH_END;
      } else if (why < 0) {
        flag = why;
        t = u.tn;
        yout = b.zn[0];
        phase = PH_END;
      }
    }
    if (phase == PH_END) {
      // ---- calculate_abundances.c retry / abort policy, next output time
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
      if (t < tout && nfails < 5) {
        phase = PH_CALL;
      } else {
        // write-back with clip (calculate_abundances.c:308-317)
        double v = yout;
        if (v < c.atol) v = 0.0;
        if (c.row_ok) ab[c.i] = v;
        const double xe_fin = a.xe_policy ? half_sum(c.row_ok ? c.q * v : 0.0, c.hh) : 0.0;
        wsync();
        if (c.lead()) {
          ab[a.nelect0] = a.xe_policy ? xe_fin : u.xe_last;
          if (a.stats) {
            int *st = a.stats + (size_t)sys * 8;
            st[0] = u.nst; st[1] = u.nfe; st[2] = u.nje; st[3] = u.nsetups;
            st[4] = u.nni; st[5] = u.ncfn; st[6] = u.netf; st[7] = flag;
          }
          a.status[sys] = status;
        }
        wsync();
        phase = PH_FETCH;
      }
    }
  }
}
