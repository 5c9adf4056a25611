// Batched variable-order BDF, one stiff system per wavefront (wave64), for
// networks of up to 32 equations (the reduced network).
//
// Included once per network inside a namespace that defines `namespace NET`
// (a generated csrc/gen/net_<name>.h).  Replaces, for one grid point per
// wavefront, the reference's per-point CVODE solve (calculate_abundances.c:
// 132-326) and the generated RHS (odes_*.c:26-148).  The integrator is the
// CVODE restatement of bdf_core.h, run with the exact Jacobian; this file
// supplies the wavefront as its team.
//
// Linear solve: unlike CVODE's dense LU, the iteration matrix is scaled by
// the error weights (W M W^-1) and inverted explicitly by Gauss-Jordan with
// partial pivoting in registers; each solve is one matrix-vector product.
// With atol = 1e-30 the plain LU solve's rounding error in trace species
// stalls the Newton iteration (1e5 steps and more per point); the scaled
// explicit inverse removes those spurious failures (~2.6e3 steps) without
// changing the converged solution.  MT (the matrix type of J and of the
// inverse) is double for the plain FP64 form and float for the mixed form;
// the residual, Newton update, error test and all n-vectors stay FP64.
//
// Layout, lane L of the wave: row i = L & 31, half hh = L >> 5.  The lane
// holds columns hh*16 .. hh*16+15 of row i of J and of the inverse; n-vectors
// are held as the lane's row value (duplicated in both halves).  Step-control
// scalars are wave-uniform and kept in the wave's LDS record.

#include "wave_ops.h"

using namespace pdrw;

#include "bdf_core.h"

// Systems (wavefronts) per block, NW, is a template parameter chosen per
// launch by the autotuner (1, 2 or 4; the per-block RHS tables are shared by
// the block's wavefronts).

// Occupancy: the kernel is latency-bound (every step is a chain of
// dependent cross-lane reductions and LDS round trips), so throughput grows
// with the wavefronts per SIMD that hide each other's latency, as long as
// the register budget that buys them does not push the integrator's state
// into scratch.  Both forms are compiled for a 168-register budget (3
// wavefronts per SIMD) and use 48.5 KB of LDS per block (3 blocks per CU).
#ifndef PDR_WPE32
#define PDR_WPE32 3
#endif
#ifndef PDR_WPE64
#define PDR_WPE64 3
#endif
template <typename MT>
constexpr int kWavesPerEU = sizeof(MT) == 4 ? PDR_WPE32 : PDR_WPE64;

// The slot rate constants (read by every RHS) stay in LDS in the FP32 form;
// the FP64 form keeps them in the device workspace (L2) so that its LDS
// (flux partials in FP64) fits three blocks per CU.
template <typename MT>
constexpr bool kKpInLDS = sizeof(MT) == 4;

// Per-wavefront workspace in device memory for the saved Jacobian, read
// only when the iteration matrix is formed without a new Jacobian.  Indexed
// [k*64 + lane] and only ever read back by the lane that wrote it, so it
// needs no synchronisation.
template <typename MT>
struct WaveWS {
  MT J[16 * 64];                // the lane's 16 columns
  double Kp[kKpInLDS<MT> ? 1 : NET::ROUNDS * 64];  // FP64 form: slot rate constants
};

// LDS words (doubles) of the flux region: FP64 fluxes F[0..NSLOT) and the
// zero F[NSLOT] for the RHS; partials D[0..2*NSLOT) in the matrix type and
// the zero D[2*NSLOT] for the Jacobian, over the same bytes
template <typename MT>
constexpr int kFWords = (NET::NSLOT + 1) > (int)(((2 * NET::NSLOT + 1) * sizeof(MT) + 7) / 8)
                            ? (NET::NSLOT + 1)
                            : (int)(((2 * NET::NSLOT + 1) * sizeof(MT) + 7) / 8);

// per-wavefront shared memory
template <typename MT>
struct WaveLDS {
  double X[64];                    // factor vector: y, x_e, 1
  double F[kFWords<MT>];           // fluxes (RHS) or flux partials (Jacobian)
  double Kp[kKpInLDS<MT> ? NET::ROUNDS * 64 : 1];  // rate * (n_H) of each lane's flux slots
  double Zh[(QMAX - 1) * 64];      // Nordsieck entries zn[2..QMAX] of each lane
  // used by one phase at a time: the RHS (P), the M build and the solve
  // (V, V2), the Gauss-Jordan steps (PR)
  union {
    struct {
      double V[64];                // vector exchange
      double V2[64];
    };
    double P[64];                  // RHS lane partials
    alignas(16) MT PR[32];         // pivot row of the Gauss-Jordan step
  };
  BdfU u;
};

// per-block shared memory: the RHS term table and the slot factors are the
// same for every wave
template <typename MT, int NW>
struct BlockLDS {
  uint32_t rterm[NET::TM * 64];
  uint32_t fab[NET::ROUNDS * 64];
  WaveLDS<MT> w[NW];
};

struct Lane {
  int L, i, hh;
  bool row_ok;      // i < NEQ
  bool owner;       // hh == 0 && row_ok: the lane that publishes row values
};

struct Tables {
  int rbase, rnl;
  double q;          // charge of row i
  const uint32_t LDSA *rterm;  // block LDS copies
  const uint32_t LDSA *fab;
};

// per-lane integrator vectors and matrices (registers)
template <typename MT>
struct Bdf {
  double zn[QMAX + 1];
  double ewt, acor, y, ftemp, wsc;
  MT A[16];
  int invp;
  // zn[2..QMAX] live in LDS between the step's bookkeeping phases (the
  // Newton loop needs only zn[0], zn[1]): registers the compiler would
  // otherwise spill to scratch around every step
  double LDSA *zh;  // zh[(j - 2) * 64] = zn[j] of this lane
  __device__ __forceinline__ void hist_in() {
#pragma unroll
    for (int j = 2; j <= QMAX; j++) zn[j] = zh[(j - 2) * 64];
  }
  __device__ __forceinline__ void hist_out() {
#pragma unroll
    for (int j = 2; j <= QMAX; j++) zh[(j - 2) * 64] = zn[j];
  }
};

// the wavefront as the integrator's team (see bdf_core.h)
template <typename MT>
struct Ctx {
  Lane ln;
  Tables tb;
  WaveLDS<MT> LDSA *S;
  WaveWS<MT> *W;
  double rtol, atol;

  __device__ __forceinline__ bool lead() const { return ln.L == 0; }
  __device__ __forceinline__ bool own() const { return ln.owner; }
  __device__ __forceinline__ double sum(double v) { return wave_sum(v); }
  // sum of non-negative values held by the owner lanes (0..31) only
  __device__ __forceinline__ double sum_own_sq(double v) { return wave_sum_lo(v); }
  __device__ __forceinline__ double vmax(double v) { return wave_max(v); }
  __device__ __forceinline__ void sync() { wsync(); }
  __device__ __forceinline__ double rhs(double y, double &xe);
  __device__ __forceinline__ int setup(BdfU LDSA &u, Bdf<MT> &b, int convfail);
  __device__ __forceinline__ double solve(const BdfU LDSA &u, const Bdf<MT> &b, double r);
};

template <typename MT>
__device__ __forceinline__ double kp_get(const Ctx<MT> &c, int s) {
  if constexpr (kKpInLDS<MT>) return c.S->Kp[s];
  else return c.W->Kp[s];
}

// ---------------------------------------------------------------------
// RHS: returns ydot_i for this lane's row; x_e (uniform) via xe_out
// ---------------------------------------------------------------------
template <typename MT>
__device__ __forceinline__ double wave_rhs(Ctx<MT> &c, double y, double &xe_out) {
  const Lane &ln = c.ln;
  WaveLDS<MT> LDSA &S = *c.S;
  if (ln.owner) S.X[ln.i] = y;
  double xe = wave_sum(ln.owner ? c.tb.q * y : 0.0);
  xe_out = xe;
  if (ln.L == 0) S.X[NET::NEQ] = xe;
  wsync();
  // all factor gathers of the flux rounds are issued before any product
  uint32_t fv[NET::ROUNDS];
  double xa[NET::ROUNDS], xb[NET::ROUNDS];
#pragma unroll
  for (int r = 0; r < NET::ROUNDS; r++) fv[r] = c.tb.fab[r * 64 + ln.L];
#pragma unroll
  for (int r = 0; r < NET::ROUNDS; r++) {
    xa[r] = S.X[fv[r] & 0xff];
    xb[r] = S.X[(fv[r] >> 8) & 0xff];
  }
#pragma unroll
  for (int r = 0; r < NET::ROUNDS; r++) S.F[r * 64 + ln.L] = kp_get(c, r * 64 + ln.L) * xa[r] * xb[r];
  wsync();
  // term word: byte offset of the flux | the coefficient's FP64 high 16
  // bits (its other bits are zero), so decoding is two masks
  // terms in groups of G: the group's words, then its gathers, then the
  // products (keeps G gathers in flight)
  constexpr int G = 8;
  double acc = 0.0;
  const unsigned char LDSA *Fb = (const unsigned char LDSA *)S.F;
#pragma unroll
  for (int t0 = 0; t0 < NET::TM; t0 += G) {
    uint32_t w[G];
    double f[G];
#pragma unroll
    for (int u = 0; u < G; u++)
      if (t0 + u < NET::TM) w[u] = c.tb.rterm[(t0 + u) * 64 + ln.L];
#pragma unroll
    for (int u = 0; u < G; u++)
      if (t0 + u < NET::TM) f[u] = *(const double LDSA *)(Fb + (w[u] & 0xffff));
#pragma unroll
    for (int u = 0; u < G; u++)
      if (t0 + u < NET::TM) acc = fma(__hiloint2double((int)(w[u] & 0xffff0000u), 0), f[u], acc);
  }
  S.P[ln.L] = acc;
  wsync();
  double yd = 0.0;
#pragma unroll
  for (int u = 0; u < NET::PM; u++) yd += (u < c.tb.rnl) ? S.P[min(c.tb.rbase + u, 63)] : 0.0;
  wsync();
  return yd;
}

// ---------------------------------------------------------------------
// exact Jacobian at the factor vector currently in S.X.  jac_partials
// writes the flux partials, jac_accumulate runs the generator's term lists:
// acc[j] = d ydot_i / d x_{hh*16+j} at fixed x_e (j < 16) and acc[16] =
// d ydot_i / d x_e, for row i of this lane; jacobian folds in the x_e chain
// term g_i * charge_j and saves the lane's 16 columns.
// ---------------------------------------------------------------------
template <typename MT>
__device__ __forceinline__ void jac_partials(Ctx<MT> &c) {
  const Lane &ln = c.ln;
  WaveLDS<MT> LDSA &S = *c.S;
  MT LDSA *D = (MT LDSA *)S.F;  // partials in the matrix type
#pragma unroll
  for (int r = 0; r < NET::ROUNDS; r++) {
    const int s = r * 64 + ln.L;
    const uint32_t f = c.tb.fab[s];
    const double kp = kp_get(c, s);
    D[2 * s] = (MT)(kp * S.X[(f >> 8) & 0xff]);   // d flux / d X[a]
    D[2 * s + 1] = (MT)(kp * S.X[f & 0xff]);      // d flux / d X[b]
  }
}

template <typename MT>
__device__ __forceinline__ void jac_accumulate(Ctx<MT> &c, double (&acc)[17]) {
  const Lane &ln = c.ln;
  const MT LDSA *D = (const MT LDSA *)c.S->F;
  // the term table lives in global memory (L2): its rounds are fetched in
  // groups of G ahead of use so that a lane waits for the table once per
  // group, not once per round; the column of each round is a compile-time
  // constant after unrolling
  constexpr int G = 8;
#pragma unroll
  for (int j = 0; j < 17; j++) acc[j] = 0.0;
  const uint32_t *jt = NET::jac_term + ln.L;
#pragma unroll
  for (int g = 0; g < NET::JT_TOTAL; g += G) {
    uint32_t w[G];
#pragma unroll
    for (int u = 0; u < G; u++)
      if (g + u < NET::JT_TOTAL) w[u] = jt[(g + u) * 64];
    MT dv[G];
#pragma unroll
    for (int u = 0; u < G; u++)
      if (g + u < NET::JT_TOTAL) dv[u] = D[w[u] & 0xffff];
#pragma unroll
    for (int u = 0; u < G; u++) {
      if (g + u < NET::JT_TOTAL) {
        const int j = NET::jac_col[g + u];
        acc[j] = fma((double)(int16_t)(w[u] >> 16), (double)dv[u], acc[j]);
      }
    }
  }
}

template <typename MT>
__device__ __forceinline__ void jacobian(Ctx<MT> &c, MT (&jv)[16]) {
  const Lane &ln = c.ln;
  WaveLDS<MT> LDSA &S = *c.S;
  jac_partials(c);
  wsync();
  double acc[17];
  jac_accumulate(c, acc);
#pragma unroll
  for (int j = 0; j < 16; j++) {
    jv[j] = (MT)fma(acc[16], (double)NET::charge[ln.hh * 16 + j], acc[j]);
    c.W->J[j * 64 + ln.L] = jv[j];
  }
  wsync();
  // in the FP64 form the partials covered the RHS's zero flux F[NSLOT]
  if (ln.L == 0) S.F[NET::NSLOT] = 0.0;
}

// ---------------------------------------------------------------------
// Gauss-Jordan inverse (in place, partial pivoting) of the scaled matrix.
// Step k: the pivot is the unused row with the largest |a[i][k]| (one DPP
// max over packed value/row keys in the half that holds column k), the
// pivot row is broadcast through LDS, the column entry of each row reaches
// the other half of the row with v_permlane32_swap.  Rows are not moved: on return A holds
// the inverse in pivot order and invp is the step at which this row was the
// pivot.  Returns false if singular.
// ---------------------------------------------------------------------
template <typename MT>
__device__ __forceinline__ bool gj_invert(const Lane &ln, WaveLDS<MT> LDSA &S, MT (&A)[16], int &invp) {
  bool used = !ln.row_ok;
  invp = ln.i;
  MT LDSA *prow = S.PR + 16 * ln.hh;  // this half's columns of the pivot row
#pragma unroll
  for (int k = 0; k < NET::NEQ; k++) {
    const int hk = k >> 4, jk = k & 15;
    const MT colk = A[jk];
    const uint32_t key = (ln.hh == hk && !used) ? pivot_key<5>(fabs(colk), ln.i) : 0u;
    bool ok;
    const int p = pivot_row<5, MT>(half_max(key, hk), ok);
    if (!ok) return false;
    const bool isp = (ln.i == p);
    // the pivot row goes through LDS: lanes p and p + 32 write their 16
    // columns, every lane reads back those of its half (one broadcast
    // address per column)
    if (isp) {
#pragma unroll
      for (int j = 0; j < 16; j++) prow[j] = A[j];
    }
    wsync();
    MT pr[16];
#pragma unroll
    for (int j = 0; j < 16; j++) pr[j] = prow[j];
    const MT piv = S.PR[k];
    wsync();
    // FP32: the hardware reciprocal (1 ulp) in place of an exactly rounded
    // division (~10 instructions); the FP32 inverse only preconditions the
    // Newton iteration
    MT d;
    if constexpr (sizeof(MT) == 4) d = __builtin_amdgcn_rcpf(piv);
    else d = (MT)1.0 / piv;
    const MT oc = other_half(colk);
    const MT fi = (ln.hh == hk) ? colk : oc;
    used = used || isp;
    invp = isp ? k : invp;
    // row update a_j <- a_j - f_i (p_j / piv), pivot row p_j / piv, column
    // k <- -f_i / piv (pivot: 1 / piv): with g the row's multiplier (-f_i/piv,
    // or 1/piv on the pivot row, which is also the new column-k entry) every
    // entry is one FMA on g and the pivot row
    const MT g = isp ? d : -fi * d;
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const MT nv = fma(g, pr[j], isp ? (MT)0 : A[j]);
      A[j] = (j == jk && ln.hh == hk) ? g : nv;
    }
  }
  return true;
}

// x = M^-1 b with the pivot-ordered in-place inverse
template <typename MT>
__device__ __forceinline__ double inv_apply(const Lane &ln, WaveLDS<MT> LDSA &S, const MT (&A)[16], int invp, double b) {
  if (ln.hh == 0) S.V[invp] = ln.row_ok ? b : 0.0;
  wsync();
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < 16; j++) acc = fma((double)A[j], S.V[ln.hh * 16 + j], acc);
  acc += other_half(acc);
  if (ln.hh == 0) S.V2[invp] = acc;
  wsync();
  double x = S.V2[ln.i];
  wsync();
  return ln.row_ok ? x : 0.0;
}

// linear-solver setup: Jacobian refresh decision, scaled Gauss-Jordan
template <typename MT>
__device__ __forceinline__ int lsetup(Ctx<MT> &c, BdfU LDSA &u, Bdf<MT> &b, int convfail) {
  const Lane &ln = c.ln;
  WaveLDS<MT> LDSA &S = *c.S;
  // b.A (the old inverse, no longer needed) receives J: the saved one,
  // fetched first so that its latency overlaps the refresh decision, or a
  // new one
#pragma unroll
  for (int j = 0; j < 16; j++) b.A[j] = c.W->J[j * 64 + ln.L];
  double dgamma = fabs(u.gamma / u.gammap - 1.0);
  bool jbad = (u.nst == 0) || (u.nst > u.nstlj + MSBJ) ||
              ((convfail == CV_FAIL_BAD_J) && !(dgamma >= DGMAX_J)) || (convfail == CV_FAIL_OTHER);
  if (jbad) {
    const int nst = u.nst;
    if (ln.L == 0) {
      u.nje++;
      u.nstlj = nst;
      u.jcur = 1;
    }
    // factor vector at the predicted state (zn[0])
    if (ln.owner) S.X[ln.i] = b.zn[0];
    double xe = wave_sum(ln.owner ? c.tb.q * b.zn[0] : 0.0);
    if (ln.L == 0) S.X[NET::NEQ] = xe;
    wsync();
    jacobian(c, b.A);
  } else {
    if (ln.L == 0) u.jcur = 0;
  }
  const double gamma = u.gamma;
  // M = I - gamma J, scaled: W M W^-1 (w = error weights)
  b.wsc = ln.row_ok ? b.ewt : 1.0;
  if (ln.hh == 0) S.V[ln.i] = 1.0 / b.wsc;
  wsync();
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int col = ln.hh * 16 + j;
    const double wjinv = S.V[col];
    double m = -gamma * (double)b.A[j];
    if (col == ln.i) m += 1.0;
    if (!ln.row_ok) m = (col == ln.i) ? 1.0 : 0.0;
    b.A[j] = (MT)(m * b.wsc * wjinv);
  }
  wsync();
  bool ok = gj_invert(ln, S, b.A, b.invp);
  return ok ? 0 : 1;
}

// solve M delta = r with the scaled explicit inverse, then CVODE's
// 2/(1+gamrat) correction for a changed gamma
template <typename MT>
__device__ __forceinline__ double lsolve(Ctx<MT> &c, const BdfU LDSA &u, const Bdf<MT> &b, double r) {
  const double gamrat = u.gamrat;
  double x = inv_apply(c.ln, *c.S, b.A, b.invp, r * b.wsc) / b.wsc;
  if (gamrat != 1.0) x *= 2.0 / (1.0 + gamrat);
  return c.ln.row_ok ? x : 0.0;
}

template <typename MT>
__device__ __forceinline__ double Ctx<MT>::rhs(double y, double &xe) { return wave_rhs(*this, y, xe); }
template <typename MT>
__device__ __forceinline__ int Ctx<MT>::setup(BdfU LDSA &u, Bdf<MT> &b, int convfail) {
  return lsetup(*this, u, b, convfail);
}
template <typename MT>
__device__ __forceinline__ double Ctx<MT>::solve(const BdfU LDSA &u, const Bdf<MT> &b, double r) {
  return lsolve(*this, u, b, r);
}

// rates of one point -> the lanes' flux constants (rate, times n_H for
// two-body slots)
template <typename MT>
__device__ __forceinline__ void load_kp(Ctx<MT> &c, const double *rate, double nH) {
#pragma unroll
  for (int r = 0; r < NET::ROUNDS; r++) {
    const int s = r * 64 + c.ln.L;
    const int re = NET::slot_reac[s];
    const uint32_t f = c.tb.fab[s];
    const double k = (re >= 0) ? rate[re] : 0.0;
    const double kps = ((f >> 16) & 1) ? k * nH : k;
    if constexpr (kKpInLDS<MT>) c.S->Kp[s] = kps;
    else c.W->Kp[s] = kps;
  }
}

// ---------------------------------------------------------------------
// one grid point: calculate_abundances.c:132-326 semantics
// ---------------------------------------------------------------------
template <typename MT>
__device__ __forceinline__ void solve_point(Ctx<MT> &c, const PointArgs &a, int sys) {
  Bdf<MT> b;
  const Lane &ln = c.ln;
  BdfU LDSA &u = c.S->u;
  load_kp(c, a.rate + (size_t)sys * a.rate_stride, a.density[sys]);
  double *ab = a.abundance + (size_t)sys * a.ab_stride;
  const double y0 = ln.row_ok ? ab[ln.i] : 0.0;
  const double xe0 = ab[a.nelect0];
  wsync();
  c.S->X[ln.L] = (ln.L == NET::NEQ + 1) ? 1.0 : 0.0;
  if (ln.L == 0) {
    c.S->F[NET::NSLOT] = 0.0;
    ((MT LDSA *)c.S->F)[2 * NET::NSLOT] = (MT)0;
    bdf_init_uniform(u, a.t0, xe0);
  }
  wsync();
  b.zh = c.S->Zh + ln.L;
  b.zn[0] = y0;
#pragma unroll
  for (int j = 1; j <= QMAX; j++) b.zn[j] = 0.0;
  b.hist_out();
  b.acor = 0.0; b.y = y0; b.ftemp = 0.0; b.wsc = 1.0; b.invp = ln.i;
#pragma unroll
  for (int j = 0; j < 16; j++) b.A[j] = (MT)0;
  b.ewt = 1.0 / (c.rtol * fabs(y0) + c.atol);

  double yout = y0;
  int flag = 0;
  const double tend = a.tend_pts ? a.tend_pts[sys] : a.tend;
  const int status = integrate(c, u, b, a, tend, yout, flag);

  // write-back with clip (calculate_abundances.c:308-317); the electron
  // slot holds x_e of the last RHS evaluation (:317) or, with xe_policy 1,
  // the charge sum of the abundances written back
  double v = yout;
  if (v < c.atol) v = 0.0;
  if (ln.owner) ab[ln.i] = v;
  const double xe_fin = a.xe_policy ? wave_sum(ln.owner ? c.tb.q * v : 0.0) : 0.0;
  wsync();
  if (ln.L == 0) {
    ab[a.nelect0] = a.xe_policy ? xe_fin : u.xe_last;
    if (a.stats) {
      int *st = a.stats + (size_t)sys * 8;
      st[0] = u.nst; st[1] = u.nfe; st[2] = u.nje; st[3] = u.nsetups;
      st[4] = u.nni; st[5] = u.ncfn; st[6] = u.netf; st[7] = flag;
    }
  }
  refill_release(a);
  if (ln.L == 0) __hip_atomic_store(a.status + sys, status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  wsync();
}

// a wavefront's constant state: lane roles, table pointers, LDS record
template <typename MT, int NW>
__device__ __forceinline__ void init_ctx(Ctx<MT> &c, BlockLDS<MT, NW> &blk, WaveWS<MT> *ws, double rtol, double atol) {
  c.ln.L = lane_id();
  c.ln.i = c.ln.L & 31;
  c.ln.hh = c.ln.L >> 5;
  c.ln.row_ok = c.ln.i < NET::NEQ;
  c.ln.owner = c.ln.row_ok && c.ln.hh == 0;
  c.S = (WaveLDS<MT> LDSA *)&blk.w[threadIdx.x >> 6];
  c.W = ws + (size_t)blockIdx.x * NW + (threadIdx.x >> 6);
  c.rtol = rtol;
  c.atol = atol;
  c.tb.rbase = NET::row_base[c.ln.L];
  c.tb.rnl = NET::row_nlanes[c.ln.L];
  c.tb.q = (double)NET::charge[c.ln.i];
  c.tb.rterm = (const uint32_t LDSA *)blk.rterm;
  c.tb.fab = (const uint32_t LDSA *)blk.fab;
}

// persistent kernel: every wavefront pulls grid points from a queue
template <typename MT, int NW>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<MT>)))
bdf_wave_kernel(PointArgs a, WaveWS<MT> *ws) {
  __shared__ BlockLDS<MT, NW> blk;
  for (int k = threadIdx.x; k < NET::TM * 64; k += blockDim.x) blk.rterm[k] = NET::rhs_term[k];
  for (int k = threadIdx.x; k < NET::ROUNDS * 64; k += blockDim.x) blk.fab[k] = NET::slot_fab[k];
  __syncthreads();
  Ctx<MT> c;
  init_ctx<MT, NW>(c, blk, ws, a.rtol, a.atol);
  const unsigned long long born = rt_ticks();
  for (;;) {
    unsigned int sys = 0;
    if (c.ln.L == 0) sys = a.published ? refill_take(a, born) : atomicAdd(a.counter, 1u);
    sys = (unsigned int)uni((int)sys);
    if (a.published ? sys == kNoPoint : sys >= (unsigned int)a.nsys) break;
    solve_point(c, a, (int)(a.published ? sys & a.ring_mask : sys));
  }
}

// The solver's Jacobian at given abundances, one point per block (an
// introspection kernel, not on the solve path): rows of the bordered form,
// out[(sys*NEQ + i)*(NEQ+1) + j] = d ydot_i / d x_j at fixed x_e for j < NEQ
// and d ydot_i / d x_e for j = NEQ, evaluated by the same partials and term
// lists as the solver's Jacobian (rounded to MT where the solver rounds).
// x_e is the charge sum of x, as in the solver's RHS.
template <typename MT>
__global__ void __launch_bounds__(64) wave_jac_kernel(const double *x, const double *rate, const double *dens,
                                                      int ab_stride, int rate_stride, double *out, WaveWS<MT> *ws) {
  __shared__ BlockLDS<MT, 1> blk;
  for (int k = threadIdx.x; k < NET::TM * 64; k += blockDim.x) blk.rterm[k] = NET::rhs_term[k];
  for (int k = threadIdx.x; k < NET::ROUNDS * 64; k += blockDim.x) blk.fab[k] = NET::slot_fab[k];
  __syncthreads();
  Ctx<MT> c;
  init_ctx<MT, 1>(c, blk, ws, 0.0, 0.0);
  const int sys = blockIdx.x;
  load_kp(c, rate + (size_t)sys * rate_stride, dens[sys]);
  const double y = c.ln.row_ok ? x[(size_t)sys * ab_stride + c.ln.i] : 0.0;
  wsync();
  c.S->X[c.ln.L] = (c.ln.L == NET::NEQ + 1) ? 1.0 : 0.0;
  wsync();
  if (c.ln.owner) c.S->X[c.ln.i] = y;
  const double xe = wave_sum(c.ln.owner ? c.tb.q * y : 0.0);
  if (c.ln.L == 0) c.S->X[NET::NEQ] = xe;
  wsync();
  jac_partials(c);
  wsync();
  double acc[17];
  jac_accumulate(c, acc);
  if (c.ln.row_ok) {
    double *row = out + ((size_t)sys * NET::NEQ + c.ln.i) * (NET::NEQ + 1);
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const int col = c.ln.hh * 16 + j;
      if (col < NET::NEQ) row[col] = (double)(MT)acc[j];
    }
    if (c.ln.hh == 0) row[NET::NEQ] = (double)(MT)acc[16];
  }
}
