// Shared device helpers for the CDNA4 (gfx950) Krylov engine.
//
// Wave64 reductions use DPP (quad_perm xor-1/xor-2, row_ror 4/8) inside each
// 16-lane row and four v_readlane reads across rows: a fixed order, so a
// reduction gives the same bits every run.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cola_krylov.h"

#define CK_WAVE 64
#define CK_BLOCK 256

#define CK_CHECK_LAUNCH()                                   \
  do {                                                      \
    hipError_t e_ = hipGetLastError();                      \
    if (e_ != hipSuccess) return CK_ERR_HIP;                \
  } while (0)

namespace ck {

// ------------------------------------------------------------ DPP reductions
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  int2 p = *reinterpret_cast<int2*>(&v);
  p.x = __builtin_amdgcn_update_dpp(0, p.x, CTRL, 0xF, 0xF, false);
  p.y = __builtin_amdgcn_update_dpp(0, p.y, CTRL, 0xF, 0xF, false);
  return *reinterpret_cast<double*>(&p);
}

__device__ __forceinline__ float readlane_f32(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  int2 p = *reinterpret_cast<int2*>(&v);
  p.x = __builtin_amdgcn_readlane(p.x, lane);
  p.y = __builtin_amdgcn_readlane(p.y, lane);
  return *reinterpret_cast<double*>(&p);
}

// Sum over the 64 lanes; every lane receives the total.
__device__ __forceinline__ float wave_sum(float v) {
  v += dpp_f32<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f32<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f32<0x124>(v);  // row_ror:4
  v += dpp_f32<0x128>(v);  // row_ror:8
  return (readlane_f32(v, 0) + readlane_f32(v, 16)) + (readlane_f32(v, 32) + readlane_f32(v, 48));
}

__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x124>(v);
  v += dpp_f64<0x128>(v);
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

__device__ __forceinline__ float wave_max(float v) {
  v = fmaxf(v, dpp_f32<0xB1>(v));
  v = fmaxf(v, dpp_f32<0x4E>(v));
  v = fmaxf(v, dpp_f32<0x124>(v));
  v = fmaxf(v, dpp_f32<0x128>(v));
  return fmaxf(fmaxf(readlane_f32(v, 0), readlane_f32(v, 16)), fmaxf(readlane_f32(v, 32), readlane_f32(v, 48)));
}

// Block-wide sum for blockDim.x == NT (multiple of 64). Result valid in all threads.
template <int NT, typename T>
__device__ __forceinline__ T block_sum(T v, T* lds /* >= NT/64 */) {
  constexpr int NW = NT / CK_WAVE;
  v = wave_sum(v);
  const int w = threadIdx.x / CK_WAVE;
  if ((threadIdx.x & (CK_WAVE - 1)) == 0) lds[w] = v;
  __syncthreads();
  T s = lds[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) s += lds[i];
  __syncthreads();
  return s;
}

// ------------------------------------------------------------ counter RNG
// SplitMix64 output function (Steele, Lea, Flood 2014).
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t stream_key(uint64_t seed, uint64_t stream) {
  return mix64(seed + 0x9E3779B97F4A7C15ull * (stream + 1));
}

__host__ __device__ __forceinline__ uint64_t stream_word(uint64_t key, uint64_t row) {
  return mix64(key + 0x9E3779B97F4A7C15ull * (row + 1));
}

constexpr uint64_t kGaussSalt = 0x5851F42D4C957F2Dull;

// Standard normal from one 64-bit word (Box-Muller on two 24-bit uniforms).
__device__ __forceinline__ float gauss_from_word(uint64_t w) {
  const float u1 = ((float)(uint32_t)(w >> 40) + 0.5f) * (1.0f / 16777216.0f);
  const float u2 = (float)(uint32_t)((w >> 16) & 0xFFFFFFull) * (1.0f / 16777216.0f);
  return sqrtf(-2.0f * logf(u1)) * cosf(6.283185307179586f * u2);
}

__host__ __device__ __forceinline__ int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace ck
