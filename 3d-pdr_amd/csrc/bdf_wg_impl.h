// Batched variable-order BDF, one stiff system per workgroup of 256 lanes
// (four wavefronts), for networks of more than 32 equations (the full
// network, 214 equations, and MYNETWORK, 57).
//
// Included once per network inside a namespace that defines `namespace NET`
// (a generated csrc/gen/wg_<name>.h).  Replaces, for one grid point per
// workgroup, the reference's per-point CVODE solve (calculate_abundances.c:
// 132-326), the generated RHS (odes_full.c:26-700) and CVODE's dense LU.
// The integrator is the CVODE restatement of bdf_core.h; this file supplies
// the workgroup as its team.
//
// Linear algebra (tables and their derivation: pdrchem/sparse.py and
// pdrchem/wggen.py).  The Newton matrix is taken in bordered form, with x_e
// as unknown NEQ, so that it stays as sparse as the network:
//     [ I - gamma J0   -gamma g ] [dx ]   [r]
//     [   -q^T            1     ] [dxe] = [0]
// scaled by the error weights (row i by w_i, column j by 1/w_j).  A static
// elimination order splits it into a head (pivots eliminated level by level
// without pivoting, multipliers in place) and a dense tail of NT <= 64
// unknowns.  The setup also forms the explicit inverses of the head's
// triangular factors, level by level.  The tail's Schur complement
// Tl - L_TH U_HT is formed with matrix-core (MFMA) products, one 32x32 tile
// per wavefront, and inverted by Gauss-Jordan with partial pivoting spread
// over the four wavefronts (each keeps 16 columns of the inverse in
// registers).  A solve is four chunked sparse mat-vecs with no levels: the
// head's L^-1, the tail rows' gather of head values, the tail inverse, the
// head rows' scatter of tail values and the head's U^-1.  MT is the matrix
// type (double: plain FP64 form; float: mixed form); the RHS, Newton
// iteration, error test and every n-vector stay FP64.
//
// Memory: n-vectors are held one row per lane (lane i < NB); step-control
// scalars in the workgroup's LDS record.  LDS holds the factor vector X, the
// solve vector B, the weights W, chunk partials P and one region that is, by
// phase: the fluxes F (RHS), the flux partials D (Jacobian evaluation), and
// the head entries Hv followed by the tail block Tl (factorisation and
// solves; Tl lies over F, which is dead while the factorisation runs).  The
// Jacobian of the last evaluation (CVODE keeps it for later setups) and the
// lanes' slot rate constants live in a per-workgroup slot of device memory.

#include "wave_ops.h"

using namespace pdrw;

#include "bdf_core.h"

constexpr int BLOCK = NET::BLOCK;
constexpr int NWAVE = BLOCK / 64;
static_assert(BLOCK == 256, "the Schur tiles assume four wavefronts");
static_assert(NET::NB <= BLOCK, "one row per lane");
static_assert(NET::NT <= 64, "the tail is one wavefront wide");

// LDS region sizes: Hv[NHV] and the zero slot (padded to TBASE entries),
// then max(F[NSLOT+1] FP64, Tl[NT*NT]); D[2*NSLOT+1] FP64 over all of it
template <typename MT>
struct RegionBytes {
  static constexpr size_t hvp = sizeof(MT) * NET::TBASE;
  static constexpr size_t f = sizeof(double) * (NET::NSLOT + 1);
  static constexpr size_t tl = sizeof(MT) * 64 * (64 + 1);  // tail block or inverse, stride 65
  static constexpr size_t d = sizeof(double) * (2 * NET::NSLOT + 1);
  static constexpr size_t ftl = f > tl ? f : tl;
  static constexpr size_t total = (hvp + ftl > d) ? hvp + ftl : d;
  static_assert(NET::TBASE % 4 == 0, "F and Tl start 16-byte aligned");
};

// Gauss-Jordan pivot of one step: one LDS read gives a wavefront both
template <typename MT>
struct alignas(2 * sizeof(MT)) GjPivot {
  MT d;
  int p;
};

template <typename MT>
struct WgLDS {
  double X[NET::NB + 1];     // factor vector: x_0..x_{NEQ-1}, x_e, 1
  double B[NET::NB + 1];     // solve vector; B[NB] = 0
  double W[NET::NB];         // error weights the matrix was scaled with
  double Winv[NET::NB];      //   and their reciprocals
  double P[NET::NV + 1];     // RHS chunk partials (P[NV] = 0), tail gather partials
  double V[64], V2[64];      // tail inverse exchange
  MT gjcol[2][64];           // Gauss-Jordan: pivot column (double-buffered by step)
  GjPivot<MT> gjpv[2];       //   reciprocal of the pivot and its row (-1: singular)
  double red[2][NWAVE];      // team reductions (double-buffered)
  int badpiv;                // head elimination met a zero or non-finite pivot
  BdfU u;
  // Hv[NHV] | zero | Tl[NT*NT] over F[NSLOT+1];  D[2*NSLOT+1] over all of it
  alignas(16) unsigned char region[RegionBytes<MT>::total];
};

// per-lane integrator vectors and a 16-column slice of the tail inverse
template <typename MT>
struct BdfWG {
  double zn[QMAX + 1];
  double ewt, acor, y, ftemp, wsc;
  MT A4[16];                 // inverse row = lane, columns 16*wave .. +15 (pivot order)
  int invp;                  // wavefront 0: step at which tail row `lane` was the pivot
  // the whole Nordsieck history stays in registers (moving zn[2..QMAX]
  // to the device slot measured no faster and costs traffic)
  __device__ __forceinline__ void hist_in() {}
  __device__ __forceinline__ void hist_out() {}
};

// Per-workgroup slot of the device workspace: the saved Jacobian, by LU
// position (the M build reads it in position order), and the lanes' slot rate
// constants (lane-private: written and read by the same lane, [r * BLOCK +
// lane]).
template <typename MT>
struct Slot {
  static constexpr size_t jac_bytes = ((size_t)NET::MTMP * BLOCK * sizeof(MT) + 255) / 256 * 256;
  static constexpr size_t kp_off = jac_bytes;
  static constexpr size_t bytes = kp_off + (size_t)NET::FR * BLOCK * sizeof(double);
};

// Pointers into device memory in the global address space: loads through
// them are global loads, which the hardware counts apart from LDS accesses,
// so a wait for LDS operands does not also wait for table words in flight
// (a generic pointer gives flat loads, counted with both).
template <typename T>
using gptr = const __attribute__((address_space(1))) T *;
template <typename T>
using gmut = __attribute__((address_space(1))) T *;

template <typename MT>
struct Team {
  int tid, lane, wave;
  bool own_;                 // tid < NEQ
  double q;                  // charge of row tid
  gmut<double> Kp;           // rate * (n_H) of this lane's flux slots: Kp[r * BLOCK]
  int rbase, rnl;            // RHS partial run of row tid
  int red_buf;
  WgLDS<MT> LDSA *S;
  gmut<MT> Jg;               // device-memory Jacobian slot, by LU position [MTMP * BLOCK]
  gptr<uint32_t> fab;        // slot factor table, this lane's column: fab[r * BLOCK]
  double rtol, atol;

  __device__ __forceinline__ bool lead() const { return tid == 0; }
  __device__ __forceinline__ bool own() const { return own_; }
  __device__ __forceinline__ void sync() { __syncthreads(); }
  __device__ __forceinline__ double sum(double v) {
    v = wave_sum(v);
    if (lane == 0) S->red[red_buf][wave] = v;
    __syncthreads();
    const double r = (S->red[red_buf][0] + S->red[red_buf][1]) + (S->red[red_buf][2] + S->red[red_buf][3]);
    red_buf ^= 1;
    return r;
  }
  __device__ __forceinline__ double sum_own_sq(double v) { return sum(v); }
  __device__ __forceinline__ double vmax(double v) {
    v = wave_max(v);
    if (lane == 0) S->red[red_buf][wave] = v;
    __syncthreads();
    const double r = fmax(fmax(S->red[red_buf][0], S->red[red_buf][1]), fmax(S->red[red_buf][2], S->red[red_buf][3]));
    red_buf ^= 1;
    return r;
  }
  __device__ __forceinline__ double rhs(double y, double &xe);
  __device__ __forceinline__ int setup(BdfU LDSA &u, BdfWG<MT> &b, int convfail);
  __device__ __forceinline__ double solve(const BdfU LDSA &u, const BdfWG<MT> &b, double r);
};

// Table base address made opaque at the point of use.  Without it the
// compiler hoists the per-round addresses of the lane tables (one 64-bit
// register pair per unrolled round) out of the integrator's loops, where they
// stay live across the whole solve and push the kernel into spills.  The
// result keeps the global address space.
template <typename T>
__device__ __forceinline__ gptr<T> opaque(const T *p) {
  asm volatile("" : "+s"(p));
  return (gptr<T>)p;
}

// Runs a lane's n table items (item t at tab[(off + t)*BLOCK + tid]) in
// groups of U: the group's words are loaded together, the LDS operands of
// the whole group are fetched by gather(w), then consume(w, g) runs on the
// group in order, so that a lane waits for the table and for LDS once per
// group.  Valid only where no item's consume writes what another item of the
// same call gathers (the Jacobian's terms write device memory only).
template <int U, typename W, typename G, typename C>
__device__ __forceinline__ void for_items_gc(gptr<W> tab, int off, int n, int tid, G &&gather, C &&consume) {
  using V = decltype(gather(W()));
  int t = 0;
  for (; t + U <= n; t += U) {
    W w[U];
#pragma unroll
    for (int u = 0; u < U; u++) w[u] = tab[(size_t)(off + t + u) * BLOCK + tid];
    V g[U];
#pragma unroll
    for (int u = 0; u < U; u++) g[u] = gather(w[u]);
#pragma unroll
    for (int u = 0; u < U; u++) consume(w[u], g[u]);
  }
  for (; t < n; t++) {
    const W w = tab[(size_t)(off + t) * BLOCK + tid];
    consume(w, gather(w));
  }
}

// Workgroup barrier that orders LDS only: __syncthreads' release fence also
// waits for the wavefront's outstanding global loads, which would expose the
// table loads issued ahead of the next phase at every barrier.  For phases
// that exchange data through LDS alone (the solve).
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Team sum across a barrier that orders LDS only (for phases whose table
// loads stay in flight across it)
template <typename TeamT>
__device__ __forceinline__ double sum_lds(TeamT &c, double v) {
  v = wave_sum(v);
  if (c.lane == 0) c.S->red[c.red_buf][c.wave] = v;
  lds_sync();
  const double r = (c.S->red[c.red_buf][0] + c.S->red[c.red_buf][1]) + (c.S->red[c.red_buf][2] + c.S->red[c.red_buf][3]);
  c.red_buf ^= 1;
  return r;
}

// The next U words of a lane's table runs, in flight ahead of their use
// (words of narrower tables are widened to W).
template <int U, typename W>
struct Ahead {
  W w[U];
  template <typename T>
  __device__ __forceinline__ void load(gptr<T> tab, int off, int n, int tid) {
#pragma unroll
    for (int u = 0; u < U; u++)
      if (u < n) w[u] = (W)tab[(size_t)(off + u) * BLOCK + tid];
  }
};

// A lane's run of N table items (item t at tab[(off + t)*BLOCK + tid]) in
// groups of U, software-pipelined: `a` holds the group's words on entry, the
// LDS operands of the whole group are fetched by gather(w), then the next
// group's words are loaded, then consume(w, g) runs on the group in order.
// The group after the run's last is the first of the run that follows
// (ntab, noff, NN items), so on return `a` holds its words and their latency
// overlaps this run's last group and the barrier between the two runs.  The
// lengths are compile-time constants: the group loop unrolls and every
// length test folds away, so all loads are unconditional and the compiler
// waits for exactly the group it needs while the next one stays in flight
// (a load under a run-time condition makes it wait for every load).  Valid
// only where no item's consume writes what another item of the same run
// gathers: the RHS terms write registers, the solve's chunks partial sums.
template <int U, int N, int NN, typename W, typename T, typename TN, typename G, typename C>
__device__ __forceinline__ void run_const(Ahead<U, W> &a, gptr<T> tab, int off, gptr<TN> ntab, int noff, int tid,
                                          G &&gather, C &&consume) {
  using V = decltype(gather(W()));
  if constexpr (N <= 0) {
    a.load(ntab, noff, NN, tid);
  } else {
#pragma unroll
    for (int t = 0; t < N; t += U) {
      V g[U];
      W cur[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        cur[u] = a.w[u];
        if (t + u < N) g[u] = gather(cur[u]);
      }
      if (t + U < N) a.load(tab, off + t + U, N - t - U, tid);
      else a.load(ntab, noff, NN, tid);
#pragma unroll
      for (int u = 0; u < U; u++)
        if (t + u < N) consume(cur[u], g[u]);
    }
  }
}

template <typename MT>
struct Pair {
  MT a, b;
};
template <typename MT>
struct Triple {
  MT a, b, t;
};
template <typename MT>
struct Quad {
  MT a, b, p, t;
};

// the region as LU storage (positions: [0, NHV) head entries, NHV the zero
// slot, TBASE + r*NT + c the tail block) and as the flux arrays
template <typename MT>
__device__ __forceinline__ MT LDSA *regionLU(WgLDS<MT> LDSA *S) {
  return (MT LDSA *)S->region;
}
template <typename MT>
__device__ __forceinline__ double LDSA *regionF(WgLDS<MT> LDSA *S) {
  return (double LDSA *)(S->region + RegionBytes<MT>::hvp);
}
template <typename MT>
__device__ __forceinline__ double LDSA *regionD(WgLDS<MT> LDSA *S) {
  return (double LDSA *)S->region;
}

// ---------------------------------------------------------------------
// RHS: fluxes, chunk sums, row sums.  Returns ydot of row tid (0 beyond NEQ)
// ---------------------------------------------------------------------
template <typename MT>
__device__ __forceinline__ double wg_rhs(Team<MT> &c, double y, double &xe_out) {
  WgLDS<MT> LDSA &S = *c.S;
  double LDSA *F = regionF(c.S);
  // the lane's flux factors and rate constants and its first term words do
  // not depend on the state: their loads are issued first and arrive under
  // the reduction; every barrier here orders LDS only, so none waits for them
  uint32_t fw[NET::FR];
  double kp[NET::FR];
#pragma unroll
  for (int r = 0; r < NET::FR; r++) {
    fw[r] = c.fab[r * BLOCK];
    kp[r] = c.Kp[r * BLOCK];
  }
  // the terms of the lane's VL chunks run interleaved -- item i is term
  // i / VL of chunk i % VL -- in groups of RG items, so a group's LDS reads
  // and the chunks' accumulation chains overlap; each chunk still adds its
  // terms in order
  constexpr int NI = NET::VL * NET::TM, RG = 2 * NET::VL;
  const gptr<uint32_t> rt = opaque(NET::rhs_term);
  auto term_word = [&](int i) { return rt[(size_t)((i % NET::VL) * NET::TM + i / NET::VL) * BLOCK + c.tid]; };
  // the words of the next two groups are in flight at any time
  uint32_t tw[RG], tw2[RG];
#pragma unroll
  for (int u = 0; u < RG; u++) {
    if (u < NI) tw[u] = term_word(u);
    if (RG + u < NI) tw2[u] = term_word(RG + u);
  }
  if (c.own_) S.X[c.tid] = y;
  const double xe = sum_lds(c, c.own_ ? c.q * y : 0.0);
  xe_out = xe;
  // x_e (factor index NEQ) is taken from the sum rather than from X
  if (c.tid == 0) F[NET::NSLOT] = 0.0;
#pragma unroll
  for (int r = 0; r < NET::FR; r++) {
    const int ia = fw[r] & 0xff, ib = (fw[r] >> 8) & 0xff;
    const double xa = (ia == NET::NEQ) ? xe : S.X[ia];
    const double xb = (ib == NET::NEQ) ? xe : S.X[ib];
    F[r * BLOCK + c.tid] = kp[r] * xa * xb;
  }
  lds_sync();
  // term word: byte offset of the flux | the coefficient's FP64 high 16
  // bits (its other bits are zero), so decoding is two masks
  const unsigned char LDSA *Fb = (const unsigned char LDSA *)F;
  double acc[NET::VL];
#pragma unroll
  for (int p = 0; p < NET::VL; p++) acc[p] = 0.0;
#pragma unroll
  for (int i0 = 0; i0 < NI; i0 += RG) {
    double f[RG];
    uint32_t cw[RG];
#pragma unroll
    for (int u = 0; u < RG; u++) {
      cw[u] = tw[u];
      if (i0 + u < NI) f[u] = *(const double LDSA *)(Fb + (cw[u] & 0xffff));
    }
#pragma unroll
    for (int u = 0; u < RG; u++) {
      tw[u] = tw2[u];
      if (i0 + 2 * RG + u < NI) tw2[u] = term_word(i0 + 2 * RG + u);
    }
#pragma unroll
    for (int u = 0; u < RG; u++)
      if (i0 + u < NI) {
        const int p = (i0 + u) % NET::VL;
        acc[p] = fma(__hiloint2double((int)(cw[u] & 0xffff0000u), 0), f[u], acc[p]);
      }
  }
#pragma unroll
  for (int p = 0; p < NET::VL; p++) S.P[p * BLOCK + c.tid] = acc[p];
  lds_sync();
  // the row's chunk partials in a fixed order, read 8 at a time (indices
  // clamped to P[NV] = 0) so that a row waits for LDS once per 8 partials
  double yd = 0.0;
  for (int u0 = 0; u0 < c.rnl; u0 += 8) {
    double p[8];
#pragma unroll
    for (int u = 0; u < 8; u++) p[u] = S.P[min(c.rbase + u0 + u, NET::NV)];
#pragma unroll
    for (int u = 0; u < 8; u++) yd = (u0 + u < c.rnl) ? yd + p[u] : yd;
  }
  return yd;
}

// ---------------------------------------------------------------------
// Jacobian of the bordered system at the state in X -> device slot Jg
// ---------------------------------------------------------------------
template <typename MT>
__device__ __forceinline__ void wg_jacobian(Team<MT> &c) {
  WgLDS<MT> LDSA &S = *c.S;
  double LDSA *D = regionD(c.S);
#pragma unroll
  for (int r = 0; r < NET::FR; r++) {
    const uint32_t f = c.fab[r * BLOCK];
    const int s = r * BLOCK + c.tid;
    const double kp = c.Kp[r * BLOCK];
    D[2 * s] = kp * S.X[(f >> 8) & 0xff];      // d flux / d X[a]
    D[2 * s + 1] = kp * S.X[f & 0xff];         // d flux / d X[b]
  }
  if (c.tid == 0) D[2 * NET::NSLOT] = 0.0;
  __syncthreads();
  double acc = 0.0;
  for_items_gc<8>(opaque(NET::jac_term), 0, NET::JTM, c.tid, [&](uint32_t w) { return D[w & 0x1fff]; },
                  [&](uint32_t w, double d) {
                    const int coef = (int)(w << 15) >> 28;  // signed 4-bit field at bit 13
                    acc = fma((double)coef, d, acc);
                    if ((w >> 17) & 1) {
                      c.Jg[w >> 18] = (MT)acc;
                      acc = 0.0;
                    }
                  });
  __syncthreads();
}

// ---------------------------------------------------------------------
// Schur complement Tl <- Tl - L_TH U_HT on the matrix cores
// ---------------------------------------------------------------------
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// Operands come from the schur32 / schur64 tables: per lane one 64-bit word
// of four LU positions for two k steps (A | B << 16 | A' << 32 | B' << 48),
// loaded SCHUR_AHEAD groups ahead of use (the tables carry padding words for
// the look-ahead past the end), so the MFMA chain waits on LDS reads only.  The order of the k steps is that of a plain
// loop over k, so every output entry sums its products in the same order.
__device__ __forceinline__ void schur_mfma(const Team<float> &c, float LDSA *LU) {
  // wave w: output tile rows 32*(w>>1).., columns 32*(w&1)..
  // v_mfma_f32_32x32x2f32: lane l supplies A[i = l%32][k = l/32] and
  // B[k = l/32][j = l%32]; result register r of lane l is
  // D[8*(r/4) + 4*(l/32) + r%4][l%32]
  constexpr int G = NET::SCHUR_G32, NS = NET::KH / 4, NG = NS / G, REM = NS % G;
  const int ti = c.wave >> 1, tj = c.wave & 1;
  const int li = c.lane & 31, lk = c.lane >> 5;
  float LDSA *Tl = LU + NET::TBASE;
  const gptr<uint64_t> tab = opaque(NET::schur32) + c.wave * NS * 64 + c.lane;
  uint64_t w[NET::SCHUR_AHEAD + 1][G];
#pragma unroll
  for (int a = 0; a < NET::SCHUR_AHEAD; a++)
#pragma unroll
    for (int u = 0; u < G; u++) w[a][u] = tab[(a * G + u) * 64];
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int i = 32 * ti + 8 * (r >> 2) + 4 * lk + (r & 3);
    acc[r] = Tl[i * NET::TS + 32 * tj + li];
  }
  // n words: every operand read from LDS first, then the MFMA chain (the
  // scheduling barriers keep the compiler from sinking the table loads and
  // the LDS reads into the chain, which would serialise their latencies)
  auto steps = [&](const uint64_t *v, auto n) {
    float x[4 * (decltype(n)::value > 0 ? decltype(n)::value : 1)];  // n = 0: a network without a head
#pragma unroll
    for (int u = 0; u < decltype(n)::value; u++) {
      x[4 * u] = -LU[v[u] & 0xffff];
      x[4 * u + 1] = LU[(v[u] >> 16) & 0xffff];
      x[4 * u + 2] = -LU[(v[u] >> 32) & 0xffff];
      x[4 * u + 3] = LU[v[u] >> 48];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < 2 * decltype(n)::value; u++)
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(x[2 * u], x[2 * u + 1], acc, 0, 0, 0);
  };
  // fully unrolled, so that the buffers rotate by name: a register copy of a
  // word still in flight would wait for its load
  constexpr int NB = NET::SCHUR_AHEAD + 1;
#pragma unroll
  for (int g = 0; g < NG; g++) {
#pragma unroll
    for (int u = 0; u < G; u++) w[(g + NET::SCHUR_AHEAD) % NB][u] = tab[((g + NET::SCHUR_AHEAD) * G + u) * 64];
    __builtin_amdgcn_sched_barrier(0);
    steps(w[g % NB], std::integral_constant<int, G>());
    __builtin_amdgcn_sched_barrier(0);
  }
  if constexpr (REM > 0) steps(w[NG % NB], std::integral_constant<int, REM>());
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int i = 32 * ti + 8 * (r >> 2) + 4 * lk + (r & 3);
    Tl[i * NET::TS + 32 * tj + li] = acc[r];
  }
}

__device__ __forceinline__ void schur_mfma(const Team<double> &c, double LDSA *LU) {
  // v_mfma_f64_16x16x4f64: lane l supplies A[i = l%16][k = l/16] and
  // B[k = l/16][j = l%16]; result register r of lane l is D[4*r + l/16][l%16].
  // wave w computes the four 16x16 tiles of row band w, interleaved: one A
  // operand (the same in every tile's word) feeds four MFMA chains.
  constexpr int NS = (NET::KH + 7) / 8;
  const int li = c.lane & 15, lk = c.lane >> 4;
  double LDSA *Tl = LU + NET::TBASE;
  const int i0 = 16 * c.wave;
  const gptr<uint64_t> tab = opaque(NET::schur64) + c.wave * NS * 4 * 64 + c.lane;
  uint64_t w[NET::SCHUR_AHEAD + 1][4];
#pragma unroll
  for (int a = 0; a < NET::SCHUR_AHEAD; a++)
#pragma unroll
    for (int t = 0; t < 4; t++) w[a][t] = tab[(a * 4 + t) * 64];
  f64x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) acc[t][r] = Tl[(i0 + 4 * r + lk) * NET::TS + 16 * t + li];
  // fully unrolled, so that the buffers rotate by name
  constexpr int NB = NET::SCHUR_AHEAD + 1;
#pragma unroll
  for (int s = 0; s < NS; s++) {
#pragma unroll
    for (int t = 0; t < 4; t++) w[(s + NET::SCHUR_AHEAD) % NB][t] = tab[((s + NET::SCHUR_AHEAD) * 4 + t) * 64];
    __builtin_amdgcn_sched_barrier(0);
    const uint64_t *v = w[s % NB];
    const double a0 = -LU[v[0] & 0xffff], a1 = -LU[(v[0] >> 32) & 0xffff];
    double b0[4], b1[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      b0[t] = LU[(v[t] >> 16) & 0xffff];
      b1[t] = LU[v[t] >> 48];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0[t], acc[t], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1[t], acc[t], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) Tl[(i0 + 4 * r + lk) * NET::TS + 16 * t + li] = acc[t][r];
}

// ---------------------------------------------------------------------
// Gauss-Jordan inverse of the tail block with partial pivoting, spread over
// the four wavefronts: lane r of every wavefront holds row r, wavefront w
// columns 16w .. 16w+15 (A4).  Step k: the wavefront that owns column k
// picks the pivot -- the unused row with the largest |a[r][k]|, ties to the
// lower row, by a DPP max over packed value/row keys -- and publishes the
// row index, its reciprocal and the whole column in LDS (double-buffered,
// so one barrier per step); every wavefront then updates its 16 columns,
// reading the pivot row's entries with v_readlane (the pivot row is taken
// into a scalar register once per step, and the 16 reads are issued
// together).  The owner of column k+1 updates that column first and
// publishes step k+1's pivot before its other columns, so the next pivot
// search overlaps the rest of step k (each entry's arithmetic is the same in
// either order).  Rows are not moved: A4 ends as the inverse in pivot order,
// invp = step at which this row was the pivot.  Returns false if the matrix
// is singular (uniform over the workgroup).  LDS only: the barriers leave
// global loads in flight.
// ---------------------------------------------------------------------
template <typename MT>
__device__ __forceinline__ bool tail_gj(Team<MT> &c, MT (&A4)[16], int &invp) {
  WgLDS<MT> LDSA &S = *c.S;
  bool used = c.lane >= NET::NT;
  invp = c.lane;
  // the owner of column k = 16 * (its wavefront) + jk publishes step k's pivot
  // (every lane divides while the search runs; the pivot lane's quotient
  // is the pivot's reciprocal, read with every lane active -- a lane-read
  // inside the lane-0 branch could see quotients the other lanes never
  // computed, once the compiler sinks the division into that branch)
  auto publish = [&](int jk, int buf) {
    const MT colk = A4[jk];
    const MT rc = (MT)1.0 / colk;
    bool ok;
    const int p = pivot_row<6, MT>(wave_max_u(used ? 0u : pivot_key<6>(fabs(colk), c.lane)), ok);
    const MT d = rdlane(rc, p);
    S.gjcol[buf][c.lane] = colk;
    if (c.lane == 0) {
      S.gjpv[buf].d = ok ? d : (MT)0.0;
      S.gjpv[buf].p = ok ? p : -1;
    }
  };
  if (c.wave == 0) publish(0, 0);
#pragma unroll 1
  for (int wk = 0; wk < 4; wk++) {
#pragma unroll
    for (int jk = 0; jk < 16; jk++) {
      const int k = 16 * wk + jk;
      if (k >= NET::NT) break;
      const int buf = k & 1;
      lds_sync();
      // the pivot and the column in one LDS round trip; a singular step is
      // left after its (meaningless) updates, by every wavefront alike
      const MT d = S.gjpv[buf].d;
      const int p = uni(S.gjpv[buf].p);
      const MT fi = S.gjcol[buf][c.lane];
      const bool isp = (c.lane == p);
      if (isp) { used = true; invp = k; }
      MT pr[16];
#pragma unroll
      for (int j = 0; j < 16; j++) pr[j] = rdlane(A4[j], p);
      // column k+1 (wavefront wn, slot jn) first, by its owner, which then
      // publishes step k+1's pivot while the rest of step k proceeds
      const int jn = (jk + 1) & 15;  // a constant once the jk loop unrolls
      const int wn = (jk == 15) ? wk + 1 : wk;
      const bool ahead = (k + 1 < NET::NT) && (c.wave == wn);
      if (ahead) {
        const MT pj = pr[jn] * d;
        A4[jn] = isp ? pj : fma(-fi, pj, A4[jn]);
        publish(jn, (k + 1) & 1);
      }
#pragma unroll
      for (int j = 0; j < 16; j++) {
        if (j == jn && ahead) continue;
        const MT pj = pr[j] * d;
        A4[j] = isp ? pj : fma(-fi, pj, A4[j]);
      }
      // the pivot column becomes the elimination's column of the inverse
      if (c.wave == wk) A4[jk] = isp ? d : -fi * d;
      if (p < 0) return false;
    }
  }
  return true;
}

// ---------------------------------------------------------------------
// setup: Jacobian refresh decision, M build, head elimination, Schur
// complement, tail inverse.  Returns 0, or 1 if a pivot vanished.
// ---------------------------------------------------------------------
template <typename MT>
__device__ __forceinline__ int wg_setup(Team<MT> &c, BdfU LDSA &u, BdfWG<MT> &b, int convfail) {
  WgLDS<MT> LDSA &S = *c.S;
  MT LDSA *LU = regionLU(c.S);
  const double gamma = u.gamma;
  const double dgamma = fabs(gamma / u.gammap - 1.0);
  const int nst = u.nst;
  const bool jbad = (nst == 0) || (nst > u.nstlj + MSBJ) ||
                    ((convfail == CV_FAIL_BAD_J) && !(dgamma >= DGMAX_J)) || (convfail == CV_FAIL_OTHER);
  __syncthreads();
  if (c.tid == 0) {
    if (jbad) {
      u.nje++;
      u.nstlj = nst;
      u.jcur = 1;
    } else {
      u.jcur = 0;
    }
  }
  // factor vector at the predicted state; x_e for the electron weight
  if (c.own_) S.X[c.tid] = b.zn[0];
  const double xe = c.sum(c.own_ ? c.q * b.zn[0] : 0.0);
  if (c.tid == 0) S.X[NET::NEQ] = xe;
  __syncthreads();
  if (jbad) wg_jacobian(c);
  // weights: W M W^-1 (w = error weights; x_e row/column by its own weight)
  b.wsc = c.own_ ? b.ewt : 1.0;
  if (c.own_) {
    S.W[c.tid] = b.wsc;
    S.Winv[c.tid] = 1.0 / b.wsc;
  }
  if (c.tid == 0) {
    const double we = 1.0 / (c.rtol * fabs(xe) + c.atol);
    S.W[NET::NEQ] = we;
    S.Winv[NET::NEQ] = 1.0 / we;
  }
  __syncthreads();
  if (c.tid == 0) S.badpiv = 0;
  // M = I - gamma J, scaled, in LU storage order, without branches.  Position
  // word d: row | col << 8 | flags -- MF_J: the saved Jacobian holds a value
  // at the same position (a J entry, or a charge of the bordered row, which
  // wg_slot_charges put there); MF_BORD: the bordered row, coefficient -1 in
  // place of -gamma; MF_DIAG: add 1; MF_FIXED: the zero or one slot (MF_ONE),
  // unscaled; MF_SKIP: padding.  A position without a J value gets -gamma * 0
  // (a signed zero).  The table entries and J values of the next group of
  // positions are loaded while the current group is scaled and stored.
  const gptr<uint32_t> mp = opaque(NET::mpos);
  const gptr<MT> jg = (gptr<MT>)c.Jg;
  // a padding position stores into its lane's slot of the tail block's
  // padding column (itself padding, never read), so that no store, and no
  // computation sunk into one, depends on a branch
  static_assert(NET::TS == NET::NT + 1, "the tail block has a padding column");
  const int mdummy = NET::TBASE + (c.tid % NET::NT) * NET::TS + NET::NT;
  auto m_entry = [&](uint32_t d, MT j, double w, double wi) -> MT {
    const double jv = (d & NET::MF_J) ? (double)j : 0.0;
    const double jt = ((d & NET::MF_BORD) ? -1.0 : -gamma) * jv;
    const double m = (d & NET::MF_DIAG) ? jt + 1.0 : jt;
    double v = m * w * wi;
    // an opaque step: the value is computed for every position, so the
    // compiler does not turn the select below into a branch around the reads
    asm volatile("" : "+v"(v));
    return (MT)((d & NET::MF_FIXED) ? ((d & NET::MF_ONE) ? 1.0 : 0.0) : v);
  };
  {
    // the position table and the Jacobian slot are padded to whole groups, so
    // every load is unconditional (the last group's prefetch reads one group
    // past them: the table's and the slot's first group, unused)
    constexpr int U = NET::MGROUP;
    static_assert(NET::MTMP % U == 0, "M build in whole groups");
    uint32_t d[U];
    MT j[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      d[u] = mp[u * BLOCK + c.tid];
      j[u] = jg[u * BLOCK + c.tid];
    }
#pragma unroll 1
    for (int t0 = 0; t0 < NET::MTMP; t0 += U) {
      uint32_t dc[U];
      MT jc[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        dc[u] = d[u];
        jc[u] = j[u];
      }
      const int tn = (t0 + U < NET::MTMP) ? t0 + U : 0;
#pragma unroll
      for (int u = 0; u < U; u++) {
        d[u] = mp[(tn + u) * BLOCK + c.tid];
        j[u] = jg[(tn + u) * BLOCK + c.tid];
      }
      // the group's scale factors first: every LDS read ahead of the stores,
      // which the compiler cannot tell apart from them
      double wr[U], wc[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        wr[u] = S.W[dc[u] & 0xff];
        wc[u] = S.Winv[(dc[u] >> 8) & 0xff];
      }
#pragma unroll
      for (int u = 0; u < U; u++)
        LU[(dc[u] & NET::MF_SKIP) ? mdummy : (t0 + u) * BLOCK + c.tid] = m_entry(dc[u], jc[u], wr[u], wc[u]);
    }
  }
  // Head elimination and head inverses: the setup stream, phase by phase
  // (one barrier each), in whole groups of SU table rounds.  The table words
  // are loaded two groups ahead of their use (the first two before the
  // barrier), unconditionally, and a phase's last groups are followed in the
  // stream by the next phase's first, so table latency overlaps the work and
  // the barriers (which order LDS only from here on).  A word holds its four
  // LU positions as byte offsets of 4-byte entries (one instruction extracts
  // each address) and its flags in the offsets' zero low bits; a phase ends
  // after the group whose first word has bit 0 set.  Stores are branch-free:
  // an item that stores nothing writes its lane's dummy slot (the tail
  // block's padding column, which nothing reads).
  const gptr<uint64_t> st = opaque(NET::setup_stream);
  Ahead<NET::SU, uint64_t> a1, a2;
  if constexpr (NET::NSP > 0) {
    a1.load(st, 0, NET::SU, c.tid);
    a2.load(st, NET::SU, NET::SU, c.tid);
  }
  lds_sync();
  if constexpr (NET::NSP > 0) {
    static_assert(NET::NT == 64 && NET::TS == 65, "the dummy slots are the tail block's padding column");
    // LU entry at byte offset off4 of a 4-byte-entry layout (position * 4)
    auto at = [&](uint32_t off4) -> MT LDSA & {
      return *(MT LDSA *)((unsigned char LDSA *)LU + off4 * (sizeof(MT) / 4));
    };
    const uint32_t dummy4 = 4u * (NET::TBASE + (c.tid & 63) * NET::TS + 64);
    int pos = 0;  // stream round of the group in a1
    // one group: gather the operands of the words in a1, shift the prefetch
    // window, consume; true if the phase ends with this group
    auto group = [&](auto &&gather, auto &&consume) -> bool {
      uint64_t cur[NET::SU];
      decltype(gather(uint64_t())) g[NET::SU];
#pragma unroll
      for (int u = 0; u < NET::SU; u++) {
        cur[u] = a1.w[u];
        g[u] = gather(cur[u]);
      }
#pragma unroll
      for (int u = 0; u < NET::SU; u++) a1.w[u] = a2.w[u];
      a2.load(st, pos + 2 * NET::SU, NET::SU, c.tid);
      pos += NET::SU;
#pragma unroll
      for (int u = 0; u < NET::SU; u++) consume(cur[u], g[u]);
      return uni((int)((uint32_t)cur[0] & 1u)) != 0;
    };
    // Elimination phase lv: the updates of the level-lv pivots and the stores
    // of the level-(lv-1) multipliers and diagonal reciprocals.  An update
    // recomputes its multiplier m = a_ik / a_kk exactly as the store does
    // (same operands, same operation), so the values equal the stored ones.
    // Every item carries its run's target, whose old value is gathered with
    // the operands (no target is an operand within a phase), so the run's
    // last item stores without waiting for an LDS read.  FP32 form: the
    // hardware reciprocal (1 ulp) in place of an exactly rounded division;
    // the factors only precondition the FP64 Newton iteration.
#pragma unroll 1
    for (int ph = 0; ph < NET::NEP; ph++) {
      MT acc = (MT)0.0;
      bool done;
      do {
        done = group(
            [&](uint64_t w) {
              const uint32_t lo = (uint32_t)w, hi = (uint32_t)(w >> 32);
              return Quad<MT>{at(lo & 0xfffcu), at(lo >> 16), at(hi & 0xfffcu), at((hi >> 16) & 0xfffcu)};
            },
            [&](uint64_t w, Quad<MT> q) {
              const uint32_t hi = (uint32_t)(w >> 32);
              MT m;
              if constexpr (sizeof(MT) == 4) m = q.a * __builtin_amdgcn_rcpf(q.p);
              else m = q.a / q.p;
              acc = fma(m, q.b, acc);
              const bool last = (hi >> 16) & 1u, store = (hi >> 17) & 1u;
              at(last ? ((hi >> 16) & 0xfffcu) : dummy4) = store ? m : q.t - acc;
              acc = last ? (MT)0.0 : acc;
            });
      } while (!done);
      lds_sync();
    }
    // The head has no pivoting: a zero or non-finite pivot makes the setup
    // fail (CVODE's lsetup failure, as the tail's singular Gauss-Jordan step
    // does) instead of passing inf/NaN on.  Checked once per pivot on its
    // final diagonal entry and stored reciprocal (FP32: a reciprocal that is
    // not finite, as for a zero or flushed denormal pivot, is rejected).
    if (c.tid < NET::NHEAD) {
      const uint32_t pc = NET::piv_chk[c.tid];
      const MT pv = at(pc & 0xffffu), rv = at(pc >> 16);
      bool bad;
      if constexpr (sizeof(MT) == 4) bad = !(isfinite(rv) && isfinite(pv));
      else bad = !(pv != (MT)0 && isfinite(pv));
      if (bad) S.badpiv = 1;
    }
    lds_sync();
    if (S.badpiv) return 1;  // uniform: read after the barrier
    // explicit inverses of the head triangles, level by level:
    // t <- -(sum of a*b) * c; the scale c is gathered with the operands
#pragma unroll 1
    for (int ph = NET::NEP; ph < NET::NSP; ph++) {
      MT acc = (MT)0.0;
      bool done;
      do {
        done = group(
            [&](uint64_t w) {
              const uint32_t lo = (uint32_t)w, hi = (uint32_t)(w >> 32);
              return Triple<MT>{at(lo & 0xfffcu), at(lo >> 16), at(hi & 0xfffcu)};
            },
            [&](uint64_t w, Triple<MT> g3) {
              const uint32_t hi = (uint32_t)(w >> 32);
              acc = fma(g3.a, g3.b, acc);
              const bool last = (hi >> 16) & 1u;
              at(last ? ((hi >> 16) & 0xfffcu) : dummy4) = -acc * g3.t;
              acc = last ? (MT)0.0 : acc;
            });
      } while (!done);
      lds_sync();
    }
    schur_mfma(c, LU);
    lds_sync();
  }
  // tail inverse: every wavefront loads its 16 columns of the block
  const MT LDSA *Tl = LU + NET::TBASE;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int col = 16 * c.wave + j;
    MT v = (c.lane == col) ? (MT)1.0 : (MT)0.0;
    if (c.lane < NET::NT && col < NET::NT) v = Tl[c.lane * NET::TS + col];
    b.A4[j] = v;
  }
  __syncthreads();
  const bool ok = tail_gj(c, b.A4, b.invp);
  __syncthreads();
  return ok ? 0 : 1;
}

// one chunk of a sparse mat-vec per lane: P[tid] = sum of LU[pos] * B[col]
// over the lane's entries (pos | col << 14); `pre` holds the chunk's first
// words on entry and the next stage's (ntab_, nrounds) on return
template <int ROUNDS, int NROUNDS, typename MT>
__device__ __forceinline__ void spmv_partials(Team<MT> &c, Ahead<8, uint32_t> &pre, const uint32_t *tab_,
                                              const uint32_t *ntab_, const MT LDSA *LU) {
  double acc = 0.0;
  run_const<8, ROUNDS, NROUNDS>(pre, opaque(tab_), 0, opaque(ntab_), 0, c.tid,
            [&](uint32_t w) { return Pair<double>{(double)LU[w & 0x3fff], c.S->B[w >> 14]}; },
            [&](uint32_t, Pair<double> g) { acc = fma(g.a, g.b, acc); });
  c.S->P[c.tid] = acc;
}

// sum of a row's chunk partials P[base .. base+cnt), in a fixed order
// (left to right from 0).  cnt <= MAXC: all MAXC partials are read at once
// (index clamped to P[NV] = 0) and the first cnt added, so a row waits for
// LDS once instead of once per partial.
template <int MAXC, typename MT>
__device__ __forceinline__ double row_total(const WgLDS<MT> LDSA &S, int base, int cnt) {
  double p[MAXC > 0 ? MAXC : 1];
#pragma unroll
  for (int u = 0; u < MAXC; u++) p[u] = S.P[min(base + u, NET::NV)];
  double v = 0.0;
#pragma unroll
  for (int u = 0; u < MAXC; u++) v = (u < cnt) ? v + p[u] : v;
  return v;
}

// ---------------------------------------------------------------------
// solve M dx = r: returns dx of row tid (0 beyond NEQ), with CVODE's
// 2/(1+gamrat) correction for a changed gamma
// ---------------------------------------------------------------------
template <typename MT>
__device__ __forceinline__ double wg_solve(Team<MT> &c, const BdfU LDSA &u, const BdfWG<MT> &b, double r) {
  WgLDS<MT> LDSA &S = *c.S;
  const MT LDSA *LU = regionLU(c.S);
  const double gamrat = u.gamrat;
  // the lane's row indices, chunk bases and counts, loaded first: their
  // latency overlaps the first stages instead of opening the row totals
  const uint32_t sl0 = opaque(NET::solve_lane0)[c.tid], sl1 = opaque(NET::solve_lane1)[c.tid];
  const int hrow = sl0 & 0xff, trow = sl1 & 0xff;
  // each mat-vec's first table words are loaded during the phase before it
  Ahead<8, uint32_t> pre;
  if constexpr (NET::NHEAD > 0) pre.load(opaque(NET::s1_tab), 0, NET::S1_ROUNDS, c.tid);
  if (c.tid < NET::NB) S.B[c.tid] = c.own_ ? r * b.wsc : 0.0;
  if (c.tid == 0) S.B[NET::NB] = 0.0;
  lds_sync();
  if constexpr (NET::NHEAD > 0) {
    // S1: y_H = b_H + Linv b_H
    spmv_partials<NET::S1_ROUNDS, NET::S2_ROUNDS>(c, pre, NET::s1_tab, NET::s2_tab, LU);
    lds_sync();
    if (c.tid < NET::NHEAD) S.B[hrow] += row_total<NET::S1_MAXC>(S, (sl0 >> 8) & 0xff, (sl1 >> 16) & 0xf);
    lds_sync();
    // S2: the tail rows' gather L_TH y_H (combined in the tail step)
    spmv_partials<NET::S2_ROUNDS, NET::S4_ROUNDS>(c, pre, NET::s2_tab, NET::s4_tab, LU);
    lds_sync();
  }
  // tail: x_T = S^-1 b_T; the permuted right-hand side from wavefront 0,
  // 16-column partial products from every wavefront, combined in wavefront 0
  if (c.wave == 0) {
    double bt = 0.0;
    if (c.lane < NET::NT) {
      bt = S.B[trow];
      if constexpr (NET::NHEAD > 0) bt -= row_total<NET::S2_MAXC>(S, (sl1 >> 8) & 0xff, (sl1 >> 20) & 0xf);
    }
    S.V[b.invp] = bt;
  }
  lds_sync();
  {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 16; j++) acc = fma((double)b.A4[j], S.V[16 * c.wave + j], acc);
    S.P[c.wave * 64 + c.lane] = acc;
  }
  lds_sync();
  if (c.wave == 0) {
    const double acc = (S.P[c.lane] + S.P[64 + c.lane]) + (S.P[128 + c.lane] + S.P[192 + c.lane]);
    S.V2[b.invp] = acc;
    wsync();
    if (c.lane < NET::NT) S.B[trow] = S.V2[c.lane];
  }
  lds_sync();
  if constexpr (NET::NHEAD > 0) {
    // S4: z_H = y_H - U_HT x_T
    spmv_partials<NET::S4_ROUNDS, NET::S5_ROUNDS>(c, pre, NET::s4_tab, NET::s5_tab, LU);
    lds_sync();
    if (c.tid < NET::NHEAD) S.B[hrow] -= row_total<NET::S4_MAXC>(S, (sl0 >> 16) & 0xff, (sl1 >> 24) & 0xf);
    lds_sync();
    // S5: x_H = Winv z_H
    spmv_partials<NET::S5_ROUNDS, 0>(c, pre, NET::s5_tab, NET::s5_tab, LU);
    lds_sync();
    if (c.tid < NET::NHEAD) S.B[hrow] = row_total<NET::S5_MAXC>(S, sl0 >> 24, sl1 >> 28);
    lds_sync();
  }
  double x = c.own_ ? S.B[c.tid] / b.wsc : 0.0;
  if (gamrat != 1.0) x *= 2.0 / (1.0 + gamrat);
  return x;
}

template <typename MT>
__device__ __forceinline__ double Team<MT>::rhs(double y, double &xe) { return wg_rhs(*this, y, xe); }
template <typename MT>
__device__ __forceinline__ int Team<MT>::setup(BdfU LDSA &u, BdfWG<MT> &b, int convfail) {
  return wg_setup(*this, u, b, convfail);
}
template <typename MT>
__device__ __forceinline__ double Team<MT>::solve(const BdfU LDSA &u, const BdfWG<MT> &b, double r) {
  return wg_solve(*this, u, b, r);
}

// rates of one point -> the lanes' flux constants (rate, times n_H for
// two-body slots)
template <typename MT>
__device__ __forceinline__ void wg_load_kp(Team<MT> &c, const double *rate, double nH) {
#pragma unroll
  for (int r = 0; r < NET::FR; r++) {
    const int re = NET::slot_reac[r * BLOCK + c.tid];
    const double k = (re >= 0) ? rate[re] : 0.0;
    c.Kp[r * BLOCK] = ((c.fab[r * BLOCK] >> 16) & 1) ? k * nH : k;
  }
}

// ---------------------------------------------------------------------
// one grid point: calculate_abundances.c:132-326 semantics
// ---------------------------------------------------------------------
template <typename MT>
__device__ __forceinline__ void wg_solve_point(Team<MT> &c, const PointArgs &a, int sys) {
  BdfWG<MT> b;
  WgLDS<MT> LDSA &S = *c.S;
  BdfU LDSA &u = S.u;
  wg_load_kp(c, a.rate + (size_t)sys * a.rate_stride, a.density[sys]);
  double *ab = a.abundance + (size_t)sys * a.ab_stride;
  const double y0 = c.own_ ? ab[c.tid] : 0.0;
  const double xe0 = ab[a.nelect0];
  __syncthreads();
  if (c.tid <= NET::NB) S.X[c.tid] = (c.tid == NET::NEQ + 1) ? 1.0 : 0.0;
  if (c.tid == 0) {
    S.P[NET::NV] = 0.0;
    bdf_init_uniform(u, a.t0, xe0);
  }
  __syncthreads();
  b.zn[0] = y0;
#pragma unroll
  for (int j = 1; j <= QMAX; j++) b.zn[j] = 0.0;
  b.hist_out();
  b.acor = 0.0; b.y = y0; b.ftemp = 0.0; b.wsc = 1.0; b.invp = c.lane;
#pragma unroll
  for (int j = 0; j < 16; j++) b.A4[j] = (MT)0;
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
  if (c.own_) ab[c.tid] = v;
  const double xe_fin = a.xe_policy ? c.sum(c.own_ ? c.q * v : 0.0) : 0.0;
  __syncthreads();
  if (c.tid == 0) {
    ab[a.nelect0] = a.xe_policy ? xe_fin : u.xe_last;
    if (a.stats) {
      int *st = a.stats + (size_t)sys * 8;
      st[0] = u.nst; st[1] = u.nfe; st[2] = u.nje; st[3] = u.nsetups;
      st[4] = u.nni; st[5] = u.ncfn; st[6] = u.netf; st[7] = flag;
    }
  }
  if (a.published) {
    refill_release(a);
    __syncthreads();
  }
  if (c.tid == 0) {
    __hip_atomic_store(a.status + sys, status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  __syncthreads();
}

// The bordered row's entries of M (the species' charges) into the Jacobian
// slot at their LU positions, which the Jacobian never writes: the M build
// then reads every entry's value from the slot, with no dependent load of
// the charge table.  Each lane writes the positions it reads in the M build.
template <typename MT>
__device__ __forceinline__ void wg_slot_charges(Team<MT> &c) {
  for (int t = 0; t < NET::MTMP; t++) {
    const uint32_t d = NET::mpos[t * BLOCK + c.tid];
    const int row = d & 0xff, col = (d >> 8) & 0xff;
    if (row == NET::NEQ && col < NET::NEQ) c.Jg[t * BLOCK + c.tid] = (MT)NET::charge[col];
  }
}

// a workgroup's constant state: thread roles, its device slot, tables
template <typename MT>
__device__ __forceinline__ void init_team(Team<MT> &c, WgLDS<MT> &lds, unsigned char *ws, double rtol, double atol) {
  c.tid = threadIdx.x;
  c.lane = c.tid & 63;
  c.wave = c.tid >> 6;
  c.own_ = c.tid < NET::NEQ;
  c.q = (double)NET::charge[c.tid];
  c.red_buf = 0;
  c.S = (WgLDS<MT> LDSA *)&lds;
  unsigned char *slot = ws + (size_t)blockIdx.x * Slot<MT>::bytes;
  c.Jg = (gmut<MT>)slot;
  c.Kp = (gmut<double>)(slot + Slot<MT>::kp_off) + c.tid;
  c.rtol = rtol;
  c.atol = atol;
  c.fab = (gptr<uint32_t>)NET::slot_fab + c.tid;
  c.rbase = (c.tid < NET::NEQ) ? NET::row_base[c.tid] : 0;
  c.rnl = (c.tid < NET::NEQ) ? NET::row_nl[c.tid] : 0;
}

// persistent kernel: every workgroup pulls grid points from a queue; jac is
// the Jacobian workspace, NJ entries per workgroup
// FP32 matrices: two workgroups per CU fit the LDS, so the register file is
// sized for two wavefronts per SIMD
template <typename MT>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(sizeof(MT) == 4 ? 2 : 1)))
bdf_wg_kernel(PointArgs a, unsigned char *ws) {
  __shared__ WgLDS<MT> lds;
  __shared__ unsigned int next;
  Team<MT> c;
  init_team(c, lds, ws, a.rtol, a.atol);
  wg_slot_charges(c);
  const unsigned long long born = rt_ticks();
  for (;;) {
    if (c.tid == 0) next = a.published ? refill_take(a, born) : atomicAdd(a.counter, 1u);
    __syncthreads();
    const unsigned int sys = next;
    __syncthreads();
    if (a.published ? sys == kNoPoint : sys >= (unsigned int)a.nsys) break;
    wg_solve_point(c, a, (int)(a.published ? sys & a.ring_mask : sys));
  }
}

// The solver's Jacobian at given abundances, one point per workgroup (an
// introspection kernel, not on the solve path): out[(sys*NEQ + i)*(NEQ+1) +
// j] = d ydot_i / d x_j at fixed x_e for j < NEQ and d ydot_i / d x_e for
// j = NEQ, read back from the solver's saved Jacobian (in MT) through the
// iteration matrix's position table.  out must be zeroed beforehand (only
// the network's structural entries are written).  x_e is the charge sum of
// x, as in the solver's RHS.
template <typename MT>
__global__ void __launch_bounds__(BLOCK) wg_jac_kernel(const double *x, const double *rate, const double *dens,
                                                       int ab_stride, int rate_stride, double *out,
                                                       unsigned char *ws) {
  __shared__ WgLDS<MT> lds;
  Team<MT> c;
  init_team(c, lds, ws, 0.0, 0.0);
  WgLDS<MT> LDSA &S = *c.S;
  const int sys = blockIdx.x;
  wg_load_kp(c, rate + (size_t)sys * rate_stride, dens[sys]);
  const double y = c.own_ ? x[(size_t)sys * ab_stride + c.tid] : 0.0;
  if (c.tid <= NET::NB) S.X[c.tid] = (c.tid == NET::NEQ + 1) ? 1.0 : 0.0;
  __syncthreads();
  if (c.own_) S.X[c.tid] = y;
  const double xe = c.sum(c.own_ ? c.q * y : 0.0);
  if (c.tid == 0) S.X[NET::NEQ] = xe;
  __syncthreads();
  wg_jacobian(c);
  double *o = out + (size_t)sys * NET::NEQ * (NET::NEQ + 1);
  for (int t = 0; t < NET::MTM; t++) {
    const uint32_t d = NET::mpos[t * BLOCK + c.tid];
    const int row = d & 0xff, col = (d >> 8) & 0xff;
    if ((d & NET::MF_J) && row < NET::NEQ && col <= NET::NEQ)
      o[(size_t)row * (NET::NEQ + 1) + col] = (double)c.Jg[t * BLOCK + c.tid];
  }
}
