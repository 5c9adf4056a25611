// Wavefront-level helpers for gfx950 (wave64): uniform broadcasts and
// cross-lane reductions used by the one-system-per-wavefront BDF kernel.
// Reductions run on DPP within each 16-lane row and combine the four rows
// with v_readlane, so they use no LDS and return wave-uniform results.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pdrw {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// make a value wave-uniform (lives in SGPRs afterwards)
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uni(double v) {
  uint64_t b = __double_as_longlong(v);
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
  return __longlong_as_double(((uint64_t)hi << 32) | lo);
}

// value of lane `l` (uniform l), as a uniform value
__device__ __forceinline__ int rdlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float rdlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double rdlane(double v, int l) {
  uint64_t b = __double_as_longlong(v);
  uint32_t lo = __builtin_amdgcn_readlane((uint32_t)b, l);
  uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(b >> 32), l);
  return __longlong_as_double(((uint64_t)hi << 32) | lo);
}

// DPP lane permutations within a 16-lane row
constexpr int DPP_XOR1 = 0xB1;         // quad_perm [1,0,3,2]
constexpr int DPP_XOR2 = 0x4E;         // quad_perm [2,3,0,1]
constexpr int DPP_HALF_MIRROR = 0x141; // lane i <-> 7-i within each half-row
constexpr int DPP_MIRROR = 0x140;      // lane i <-> 15-i within the row

template <int CTRL>
__device__ __forceinline__ int dpp(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(dpp<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  uint64_t b = __double_as_longlong(v);
  uint32_t lo = (uint32_t)dpp<CTRL>((int)(uint32_t)b);
  uint32_t hi = (uint32_t)dpp<CTRL>((int)(uint32_t)(b >> 32));
  return __longlong_as_double(((uint64_t)hi << 32) | lo);
}

// the same lane of the other 32-lane half (v_permlane32_swap).  Swapping a
// register with itself leaves each lane holding {own value, partner value}
// in the two results (which element is which depends on the lane's half);
// the partner is the element that is not the lane's own word, and when both
// words are equal either one is the partner's value.
__device__ __forceinline__ int other_half(int v) {
  auto r = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
  return (int)((r[0] == (unsigned)v) ? r[1] : r[0]);
}
__device__ __forceinline__ float other_half(float v) { return __int_as_float(other_half(__float_as_int(v))); }
__device__ __forceinline__ double other_half(double v) {
  uint64_t b = __double_as_longlong(v);
  uint32_t lo = (uint32_t)other_half((int)(uint32_t)b);
  uint32_t hi = (uint32_t)other_half((int)(uint32_t)(b >> 32));
  return __longlong_as_double(((uint64_t)hi << 32) | lo);
}

// sum over each 16-lane row (every lane of the row receives it)
template <typename T>
__device__ __forceinline__ T row_sum(T v) {
  v += dpp<DPP_XOR1>(v);
  v += dpp<DPP_XOR2>(v);
  v += dpp<DPP_HALF_MIRROR>(v);
  v += dpp<DPP_MIRROR>(v);
  return v;
}

// sum over all 64 lanes, result uniform; fixed association order
__device__ __forceinline__ double wave_sum(double v) {
  v = row_sum(v);
  return (rdlane(v, 0) + rdlane(v, 16)) + (rdlane(v, 32) + rdlane(v, 48));
}

// the same for a value that is zero outside lanes 0..31 and non-negative
// (the squares of a norm): the two upper rows contribute +0, which leaves the
// sum unchanged, so they are not read
__device__ __forceinline__ double wave_sum_lo(double v) {
  v = row_sum(v);
  return rdlane(v, 0) + rdlane(v, 16);
}

// max over all lanes, result uniform
__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp<DPP_XOR1>(v));
  v = fmax(v, dpp<DPP_XOR2>(v));
  v = fmax(v, dpp<DPP_HALF_MIRROR>(v));
  v = fmax(v, dpp<DPP_MIRROR>(v));
  return fmax(fmax(rdlane(v, 0), rdlane(v, 16)), fmax(rdlane(v, 32), rdlane(v, 48)));
}

// Pivot search by one unsigned max.  The key of a candidate |a| >= 0 in
// row idx (< 2^IB) is the value's leading word (FP32: the whole word; FP64:
// the high word) with its low IB bits replaced by 2^IB - 1 - idx.  For
// non-negative values the key orders like the value, so the largest key
// names the largest candidate, ties going to the smaller row; values closer
// than about 2^(IB-23) relative (FP32) or 2^(IB-20) (FP64) count as ties,
// which is as good a pivot either way.  Non-candidates pass key 0.
__device__ __forceinline__ uint32_t lead_word(float a) { return __float_as_uint(a); }
__device__ __forceinline__ uint32_t lead_word(double a) { return (uint32_t)(__double_as_longlong(a) >> 32); }

template <int IB, typename T>
__device__ __forceinline__ uint32_t pivot_key(T a, int idx) {
  constexpr uint32_t M = (1u << IB) - 1;
  return (lead_word(a) & ~M) | (M - (uint32_t)idx);
}

// row of the winning key, and whether its value is a usable pivot (neither
// zero nor inf/NaN)
template <int IB, typename T>
__device__ __forceinline__ int pivot_row(uint32_t kmax, bool &ok) {
  constexpr uint32_t M = (1u << IB) - 1;
  constexpr uint32_t LIM = sizeof(T) == 4 ? 0x7f800000u : 0x7ff00000u;
  ok = (kmax > M) && (kmax < LIM);
  return (int)(M - (kmax & M));
}

// max over each 16-lane row (every lane of the row receives it)
__device__ __forceinline__ uint32_t row_max(uint32_t k) {
  k = max(k, (uint32_t)dpp<DPP_XOR1>((int)k));
  k = max(k, (uint32_t)dpp<DPP_XOR2>((int)k));
  k = max(k, (uint32_t)dpp<DPP_HALF_MIRROR>((int)k));
  k = max(k, (uint32_t)dpp<DPP_MIRROR>((int)k));
  return k;
}
__device__ __forceinline__ uint32_t rdlane_u(uint32_t v, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}

// largest key over the 32 lanes of half h (uniform h), result uniform
__device__ __forceinline__ uint32_t half_max(uint32_t k, int h) {
  k = row_max(k);
  return max(rdlane_u(k, 32 * h), rdlane_u(k, 32 * h + 16));
}
// largest key over all 64 lanes, result uniform
__device__ __forceinline__ uint32_t wave_max_u(uint32_t k) {
  k = row_max(k);
  return max(max(rdlane_u(k, 0), rdlane_u(k, 16)), max(rdlane_u(k, 32), rdlane_u(k, 48)));
}

// order LDS traffic of this wavefront (no cross-wave synchronisation: every
// wavefront of the kernel works on its own system).  LDS instructions of a
// wavefront complete in order; this only stops the compiler from moving
// LDS accesses across the point.
__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

}  // namespace pdrw
