// Wavefront-level helpers for gfx950 (wave64): uniform broadcasts and
// cross-lane reductions used by the one-system-per-wavefront BDF kernel.
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

__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 64); }

// sum over all 64 lanes, result uniform
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) v += shfl_xor_d(v, m);
  return uni(v);
}

// max over all lanes, result uniform
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) v = fmax(v, shfl_xor_d(v, m));
  return uni(v);
}

// arg-max of v over lanes (ties -> smaller idx); returns uniform value, idx
__device__ __forceinline__ void wave_argmax(double v, int idx, double &vmax, int &imax) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
    double ov = shfl_xor_d(v, m);
    int oi = __shfl_xor(idx, m, 64);
    bool take = (ov > v) || (ov == v && oi < idx);
    v = take ? ov : v;
    idx = take ? oi : idx;
  }
  vmax = uni(v);
  imax = uni(idx);
}

// order LDS traffic of this wavefront (no cross-wave synchronisation: every
// wavefront of the kernel works on its own system)
__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

}  // namespace pdrw
