// Sketch products Y = scale * A G + shift * G on BF16 MFMA (FP32 accumulation).
//
// The low-rank preconditioners start from Y = K G for a dense n x n kernel matrix K and a
// sketch of k <= 64 columns (the learned network's output, a Gaussian Nystrom sketch;
// cola/linalg/preconditioning/preconditioners.py:147 for the latter).  At k = 64 the
// product carries 64 multiply-adds per matrix element: FP32 vector code is compute bound
// there (eight 8-column launches re-stream K), while v_mfma_f32_32x32x16_bf16 keeps one
// pass over K at the HBM rate.  K is read as fp32 and rounded to bf16 in registers
// (v_cvt_pk_bf16_f32); G
// arrives already rounded (ck_round_bf16).  Rounding the sketch is free: with G~ the
// rounded sketch, Y (G~^T Y)^-1 Y^T is an exact Nystrom approximation for the sketch G~.
// Rounding K is not: the preconditioner then approximates K~ = K + E, and the preconditioned
// spectrum moves by about ||E|| / noise -- with bf16 (||E|| ~ 2^-9 sqrt(n) rms(K)) that is
// O(1) on low-noise kernel systems and costs PCG steps.  With split = 1, K = K_hi + K_lo
// (both bf16, 16 significant bits together) and each tile takes two MFMAs, which leaves
// ||E|| 256 times smaller at the same single pass over K.
//
// Work split: a workgroup owns 32 rows of A and all k output columns; its four waves take
// quarters of the reduction index, each issuing 4 MFMA k-steps (64 indices) per round while
// the next round's loads are in flight; the four partial tiles are summed through LDS.
// Operand layout of v_mfma_f32_32x32x16_bf16: lane l holds A[l % 32][8 (l / 32) + 0..7]
// and B[8 (l / 32) + 0..7][l % 32]; accumulator v holds D[8 (v / 4) + 4 (l / 32) + v % 4][l % 32].
#include "common.hpp"

using namespace ck;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kRows = 32;     // rows of A per workgroup
constexpr int kWaves = 4;     // split of the reduction index
constexpr int kUnroll = 4;    // MFMA k-steps (16 indices each) per round
constexpr int kMaxCols = 64;  // sketch columns

__device__ __forceinline__ uint32_t bf16_bits(float f) {
  // round to nearest even (finite inputs; NaN stays NaN)
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u) return (u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

// eight fp32 values as bf16 (v_cvt_pk_bf16_f32: round to nearest even), and optionally the
// rounding errors, themselves rounded to bf16: hi + lo carries 16 significant bits
__device__ __forceinline__ void split8(const float4 a, const float4 b, bf16x8& hi, bf16x8* lo) {
  const float x[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const __bf16 h = (__bf16)x[i];
    hi[i] = h;
    if (lo) (*lo)[i] = (__bf16)(x[i] - (float)h);
  }
}

__device__ __forceinline__ bf16x8 as_bf16x8(const uint4 q) {
  union {
    uint4 q;
    bf16x8 v;
  } u;
  u.q = q;
  return u.v;
}

__global__ __launch_bounds__(256) void round_bf16_kernel(int64_t n, int rows, const float* __restrict__ X,
                                                         int64_t ldx, uint16_t* __restrict__ Xb, int64_t ldb,
                                                         float* __restrict__ Xr, int64_t ldr) {
  const int j = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const uint32_t h = bf16_bits(X[(int64_t)j * ldx + i]);
    if (Xb) Xb[(int64_t)j * ldb + i] = (uint16_t)h;
    if (Xr) Xr[(int64_t)j * ldr + i] = __uint_as_float(h << 16);
  }
}

// Yt[c][i] = scale * sum_l A[i][l] Gt[c][l] + shift * Gs[c][i] for c < NC (NC = 32 or 64);
// SPLIT: A = A_hi + A_lo, two bf16 products per tile
template <int NC, bool SPLIT>
__global__ __launch_bounds__(64 * kWaves) void sketch_bf16_kernel(int n, const float* __restrict__ A, int64_t lda,
                                                                  const uint16_t* __restrict__ Gt, int64_t ldg,
                                                                  float scale, float shift,
                                                                  const float* __restrict__ Gs, int64_t ldgs,
                                                                  float* __restrict__ Yt, int64_t ldy) {
  constexpr int NT = NC / 32;  // 32-column output tiles
  __shared__ float red[kWaves][NC][kRows + 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row0 = blockIdx.x * kRows;
  const int r = lane & 31, half = lane >> 5;
  // this wave's quarter of the reduction index (n is a multiple of 256)
  const int span = n / kWaves;
  const int l_begin = wave * span, l_end = l_begin + span;
  const int arow = row0 + r < n ? row0 + r : n - 1;  // rows past n read a valid row, their output is dropped
  const float* Ar = A + (int64_t)arow * lda + 8 * half;
  const uint16_t* Gc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) Gc[t] = Gt + (int64_t)(32 * t + r) * ldg + 8 * half;
  f32x16 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[t][v] = 0.f;
  // software pipeline: the next round's loads are in flight while this round's MFMAs run
  float4 a[kUnroll][2], an[kUnroll][2];
  uint4 g[kUnroll][NT], gn[kUnroll][NT];
  auto load = [&](int l0, float4 (&aa)[kUnroll][2], uint4 (&gg)[kUnroll][NT]) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const float4* ap = reinterpret_cast<const float4*>(Ar + l0 + 16 * u);
      aa[u][0] = ap[0];
      aa[u][1] = ap[1];
#pragma unroll
      for (int t = 0; t < NT; ++t) gg[u][t] = *reinterpret_cast<const uint4*>(Gc[t] + l0 + 16 * u);
    }
  };
  // The rows of A (and of G) lie a whole row apart -- the same offset modulo a large power of
  // two -- so workgroups that walked their ranges in step would all hit the same HBM channels;
  // each workgroup starts its walk at its own rotation of the range instead.
  constexpr int kStep = 16 * kUnroll;
  const int rounds = span / kStep;
  const int rot = (int)((blockIdx.x * 37u) % (unsigned)rounds);
  auto l_of = [&](int q) { return l_begin + ((q + rot) % rounds) * kStep; };
  (void)l_end;
  load(l_of(0), an, gn);
  for (int q = 0; q < rounds; ++q) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      a[u][0] = an[u][0];
      a[u][1] = an[u][1];
#pragma unroll
      for (int t = 0; t < NT; ++t) g[u][t] = gn[u][t];
    }
    if (q + 1 < rounds) load(l_of(q + 1), an, gn);
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      bf16x8 av, al;
      split8(a[u][0], a[u][1], av, SPLIT ? &al : nullptr);
#pragma unroll
      for (int t = 0; t < NT; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, as_bf16x8(g[u][t]), acc[t], 0, 0, 0);
      if (SPLIT) {
#pragma unroll
        for (int t = 0; t < NT; ++t)
          acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, as_bf16x8(g[u][t]), acc[t], 0, 0, 0);
      }
    }
  }
  // D[i][j] = acc[v] with i = 8 (v / 4) + 4 half + v % 4, j = r: row i of A, output column j
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int v = 0; v < 16; ++v) red[wave][32 * t + r][8 * (v / 4) + 4 * half + (v % 4)] = acc[t][v];
  __syncthreads();
  for (int e = threadIdx.x; e < NC * kRows; e += 64 * kWaves) {
    const int c = e / kRows, i = e % kRows;
    const int row = row0 + i;
    if (row >= n) continue;
    const float s = (red[0][c][i] + red[1][c][i]) + (red[2][c][i] + red[3][c][i]);
    float y = scale * s;
    if (shift != 0.f) y = fmaf(shift, Gs[(int64_t)c * ldgs + row], y);
    Yt[(int64_t)c * ldy + row] = y;
  }
}

}  // namespace

extern "C" {

int ck_round_bf16(int64_t n, int32_t rows, const float* X, int64_t ldx, uint16_t* Xb, int64_t ldb, float* Xr,
                  int64_t ldr, void* stream) {
  if (n < 1 || rows < 1 || !X || (!Xb && !Xr) || ldx < n || (Xb && ldb < n) || (Xr && ldr < n)) return CK_ERR_BAD_ARG;
  const int64_t g = (n + 255) / 256;
  hipLaunchKernelGGL(round_bf16_kernel, dim3((unsigned)(g > 4096 ? 4096 : g), rows), dim3(256), 0,
                     (hipStream_t)stream, n, (int)rows, X, ldx, Xb, ldb, Xr, ldr);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_sketch_bf16_supported(int64_t n, int32_t k, int64_t lda, const void* A, int64_t ldg, const void* G) {
  auto al16 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  return n >= 256 && n % 256 == 0 && n < INT32_MAX && (k == 32 || k == 64) && lda >= n && lda % 4 == 0 &&
         ldg >= n && ldg % 8 == 0 && al16(A) && al16(G);
}

int ck_sketch_bf16(int64_t n, int32_t k, const float* A, int64_t lda, float scale, float shift, const uint16_t* Gt,
                   int64_t ldg, const float* Gs, int64_t ldgs, float* Yt, int64_t ldy, int32_t split, void* stream) {
  if (!A || !Gt || !Yt || ldy < n || (shift != 0.f && (!Gs || ldgs < n))) return CK_ERR_BAD_ARG;
  if (!ck_sketch_bf16_supported(n, k, lda, A, ldg, Gt)) return CK_ERR_UNSUPPORTED;
  const unsigned grid = (unsigned)((n + kRows - 1) / kRows);
  hipStream_t s = (hipStream_t)stream;
#define CK_SKETCH(NC_, SP_)                                                                                    \
  hipLaunchKernelGGL((sketch_bf16_kernel<NC_, SP_>), dim3(grid), dim3(64 * kWaves), 0, s, (int)n, A, lda, Gt, ldg, \
                     scale, shift, Gs, ldgs, Yt, ldy)
  if (k == 64) {
    if (split)
      CK_SKETCH(64, true);
    else
      CK_SKETCH(64, false);
  } else {
    if (split)
      CK_SKETCH(32, true);
    else
      CK_SKETCH(32, false);
  }
#undef CK_SKETCH
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // extern "C"
