// Operator kernels: y = scale * S x + shift * x on column blocks.
//
//   dense  : each wavefront owns ROWS rows; lanes stride the row in float4 chunks,
//            one DPP wave reduction per (row, rhs).            [HBM / L2 bound]
//   csr    : one lane per row, short rows batched into registers (short rows) or
//            one wavefront per row (long rows);
//            int32 column indices; coalesced rowptr/value streams. [HBM bound]
//   kron   : factors of order <= 64; the last two modes are fused per slab
//            (U = K2 S K3^T with S the 64x64 slab) and the first mode is a
//            64-row GEMM over 64-column tiles, both on v_mfma_f32_16x16x4f32 with
//            LDS-staged, bank-padded factors.                    [MFMA bound]
#include "common.hpp"
#include "ops_internal.hpp"

using namespace ck;

typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace {

// ============================================================== dense fp32
template <int NB, int ROWS>
__global__ __launch_bounds__(256) void dense_f32_kernel(int n, const float* __restrict__ A, int64_t lda,
                                                        const float* __restrict__ X, int64_t ldx, float* __restrict__ Y,
                                                        int64_t ldy, int nbr, float scale, float shift,
                                                        const int* __restrict__ done, bool vec) {
  if (done && *done) return;
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int row0 = wave * ROWS;
  if (row0 >= n) return;
  float acc[ROWS][NB];
#pragma unroll
  for (int r = 0; r < ROWS; ++r)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[r][b] = 0.f;

  if (vec) {
    for (int j = lane * 4; j < n; j += 256) {
      float4 xv[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b)
        xv[b] = (b < nbr) ? *reinterpret_cast<const float4*>(X + b * ldx + j) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        if (row0 + r < n) {
          const float4 a = *reinterpret_cast<const float4*>(A + (int64_t)(row0 + r) * lda + j);
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            acc[r][b] = fmaf(a.x, xv[b].x, acc[r][b]);
            acc[r][b] = fmaf(a.y, xv[b].y, acc[r][b]);
            acc[r][b] = fmaf(a.z, xv[b].z, acc[r][b]);
            acc[r][b] = fmaf(a.w, xv[b].w, acc[r][b]);
          }
        }
      }
    }
  } else {
    for (int j = lane; j < n; j += 64) {
      float xv[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) xv[b] = (b < nbr) ? X[b * ldx + j] : 0.f;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        if (row0 + r < n) {
          const float a = A[(int64_t)(row0 + r) * lda + j];
#pragma unroll
          for (int b = 0; b < NB; ++b) acc[r][b] = fmaf(a, xv[b], acc[r][b]);
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const float s = wave_sum(acc[r][b]);
      const int row = row0 + r;
      if (lane == 0 && b < nbr && row < n) {
        float y = scale * s;
        if (shift != 0.f) y = fmaf(shift, X[b * ldx + row], y);
        Y[b * ldy + row] = y;
      }
    }
  }
}

// bf16 stored matrix, fp32 vectors and accumulation (vector path; the MFMA
// path for nrhs >= 4 is in mixed.hip).
__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }

template <int NB, int ROWS>
__global__ __launch_bounds__(256) void dense_bf16_kernel(int n, const uint16_t* __restrict__ A, int64_t lda,
                                                         const float* __restrict__ X, int64_t ldx, float* __restrict__ Y,
                                                         int64_t ldy, int nbr, float scale, float shift,
                                                         const int* __restrict__ done) {
  if (done && *done) return;
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int row0 = wave * ROWS;
  if (row0 >= n) return;
  float acc[ROWS][NB];
#pragma unroll
  for (int r = 0; r < ROWS; ++r)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[r][b] = 0.f;
  // 8 bf16 (16 bytes) per lane per chunk; requires n % 8 == 0 and lda % 8 == 0 (checked by host)
  for (int j = lane * 8; j < n; j += 512) {
    float4 x0[NB], x1[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      if (b < nbr) {
        x0[b] = *reinterpret_cast<const float4*>(X + b * ldx + j);
        x1[b] = *reinterpret_cast<const float4*>(X + b * ldx + j + 4);
      } else {
        x0[b] = x1[b] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      if (row0 + r < n) {
        const uint4 w = *reinterpret_cast<const uint4*>(A + (int64_t)(row0 + r) * lda + j);
        const float a0 = bf16_lo(w.x), a1 = bf16_hi(w.x), a2 = bf16_lo(w.y), a3 = bf16_hi(w.y);
        const float a4 = bf16_lo(w.z), a5 = bf16_hi(w.z), a6 = bf16_lo(w.w), a7 = bf16_hi(w.w);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          float s = acc[r][b];
          s = fmaf(a0, x0[b].x, s);
          s = fmaf(a1, x0[b].y, s);
          s = fmaf(a2, x0[b].z, s);
          s = fmaf(a3, x0[b].w, s);
          s = fmaf(a4, x1[b].x, s);
          s = fmaf(a5, x1[b].y, s);
          s = fmaf(a6, x1[b].z, s);
          s = fmaf(a7, x1[b].w, s);
          acc[r][b] = s;
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const float s = wave_sum(acc[r][b]);
      const int row = row0 + r;
      if (lane == 0 && b < nbr && row < n) {
        float y = scale * s;
        if (shift != 0.f) y = fmaf(shift, X[b * ldx + row], y);
        Y[b * ldy + row] = y;
      }
    }
  }
}

// Multi-rhs dense product with the vector chunk staged in LDS: the eight waves of a
// block share one copy of X[:, c0:c0+1024) instead of each wave re-reading it from
// L2, so per block and chunk the traffic is 8*ROWS rows of A from HBM against one
// NB x 1024 slice of X.  A is streamed with non-temporal loads (it is read once per
// product and would otherwise evict X and the vectors of the surrounding solver from
// the caches); measured on MI355X at 16384^2 x 8 rhs: 6.3 TB/s with 8 waves x 4 rows
// and non-temporal loads, 4.6 TB/s with 4 waves x 8 rows and ordinary loads.
// Requires n % 4 == 0 (fp32) or n % 8 == 0 (bf16) and 16-byte aligned rows; lanes take
// 4 (fp32) or 8 (bf16) consecutive columns.
constexpr int kDenseChunk = 1024;
constexpr int kDenseWaves = 8;
typedef float f32x4_nt __attribute__((ext_vector_type(4)));
typedef float f32x2_nt __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4_nt __attribute__((ext_vector_type(4)));

// With astride != 0 the launch is batched over blockIdx.y: block row b multiplies its own
// matrix (A + b * astride elements) with its own vectors (X + b * xbs, Y + b * ybs).
template <int NB, int ROWS, bool BF>
__global__ __launch_bounds__(64 * kDenseWaves) void dense_lds_kernel(int n, const void* __restrict__ Av, int64_t lda,
                                                                     const float* __restrict__ X, int64_t ldx,
                                                                     float* __restrict__ Y, int64_t ldy, int nbr,
                                                                     float scale, float shift,
                                                                     const int* __restrict__ done, int64_t astride,
                                                                     int64_t xbs, int64_t ybs,
                                                                     float* __restrict__ xy_part = nullptr) {
  if (done && *done) return;
  {
    const int64_t pb = blockIdx.y;
    Av = BF ? static_cast<const void*>(static_cast<const uint16_t*>(Av) + pb * astride)
            : static_cast<const void*>(static_cast<const float*>(Av) + pb * astride);
    X += pb * xbs;
    Y += pb * ybs;
  }
  __shared__ float4 xs[NB][kDenseChunk / 4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row0 = blockIdx.x * (kDenseWaves * ROWS) + wave * ROWS;
  float acc[ROWS][NB];
#pragma unroll
  for (int r = 0; r < ROWS; ++r)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[r][b] = 0.f;
  for (int c0 = 0; c0 < n; c0 += kDenseChunk) {
    const int cn = n - c0 < kDenseChunk ? n - c0 : kDenseChunk;  // multiple of 4
    __syncthreads();
    for (int e = threadIdx.x; e < NB * (kDenseChunk / 4); e += 64 * kDenseWaves) {
      const int b = e / (kDenseChunk / 4), j4 = e % (kDenseChunk / 4);
      xs[b][j4] = (b < nbr && 4 * j4 < cn) ? *reinterpret_cast<const float4*>(X + (int64_t)b * ldx + c0 + 4 * j4)
                                            : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();
    if (BF) {
      const uint16_t* A = static_cast<const uint16_t*>(Av);
#pragma unroll
      for (int q = 0; q < kDenseChunk / 512; ++q) {
        const int j = q * 512 + lane * 8;
        if (j < cn) {
          uint4 w[ROWS];
#pragma unroll
          for (int r = 0; r < ROWS; ++r) {
            if (row0 + r < n) {
              const u32x4_nt v = __builtin_nontemporal_load(
                  reinterpret_cast<const u32x4_nt*>(A + (int64_t)(row0 + r) * lda + c0 + j));
              w[r] = make_uint4(v.x, v.y, v.z, v.w);
            } else {
              w[r] = make_uint4(0u, 0u, 0u, 0u);
            }
          }
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            const float4 x0 = xs[b][j / 4], x1 = xs[b][j / 4 + 1];
#pragma unroll
            for (int r = 0; r < ROWS; ++r) {
              float t = acc[r][b];
              t = fmaf(__uint_as_float(w[r].x << 16), x0.x, t);
              t = fmaf(__uint_as_float(w[r].x & 0xFFFF0000u), x0.y, t);
              t = fmaf(__uint_as_float(w[r].y << 16), x0.z, t);
              t = fmaf(__uint_as_float(w[r].y & 0xFFFF0000u), x0.w, t);
              t = fmaf(__uint_as_float(w[r].z << 16), x1.x, t);
              t = fmaf(__uint_as_float(w[r].z & 0xFFFF0000u), x1.y, t);
              t = fmaf(__uint_as_float(w[r].w << 16), x1.z, t);
              t = fmaf(__uint_as_float(w[r].w & 0xFFFF0000u), x1.w, t);
              acc[r][b] = t;
            }
          }
        }
      }
    } else {
      const float* A = static_cast<const float*>(Av);
#pragma unroll
      for (int q = 0; q < kDenseChunk / 256; ++q) {
        const int j = q * 256 + lane * 4;
        if (j < cn) {
          float4 a[ROWS];
#pragma unroll
          for (int r = 0; r < ROWS; ++r) {
            if (row0 + r < n) {
              const f32x4_nt v = __builtin_nontemporal_load(
                  reinterpret_cast<const f32x4_nt*>(A + (int64_t)(row0 + r) * lda + c0 + j));
              a[r] = make_float4(v.x, v.y, v.z, v.w);
            } else {
              a[r] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
          }
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            const float4 xv = xs[b][j / 4];
#pragma unroll
            for (int r = 0; r < ROWS; ++r) {
              float t = acc[r][b];
              t = fmaf(a[r].x, xv.x, t);
              t = fmaf(a[r].y, xv.y, t);
              t = fmaf(a[r].z, xv.z, t);
              t = fmaf(a[r].w, xv.w, t);
              acc[r][b] = t;
            }
          }
        }
      }
    }
  }
  // X . Y partials: lane L < ROWS * NB owns (row0 + L / NB, rhs L % NB) and loads its X value
  // once; the wave-wide sums give it Y
  const int my_r = lane / NB, my_b = lane % NB;
  const bool mine = xy_part && lane < ROWS * NB && my_b < nbr && row0 + my_r < n;
  const float my_x = mine ? X[(int64_t)my_b * ldx + row0 + my_r] : 0.f;
  float my_y = 0.f;
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const float sum = wave_sum(acc[r][b]);
      const int row = row0 + r;
      float y = scale * sum;
      if (lane == 0 && b < nbr && row < n) {
        if (shift != 0.f) y = fmaf(shift, X[(int64_t)b * ldx + row], y);
        Y[(int64_t)b * ldy + row] = y;
      }
      if (lane == r * NB + b) my_y = shift != 0.f ? fmaf(shift, my_x, scale * sum) : y;
    }
  }
  if (xy_part) {
    // per-block partial sums of X . Y (the CG step's p.Ap), rows then waves in index order
    __shared__ float xys[kDenseWaves][ROWS * NB];
    if (lane < ROWS * NB) xys[wave][lane] = mine ? my_x * my_y : 0.f;
    __syncthreads();
    if (threadIdx.x < nbr) {
      float t = 0.f;
      for (int w = 0; w < kDenseWaves; ++w)
#pragma unroll
        for (int r = 0; r < ROWS; ++r) t += xys[w][r * NB + threadIdx.x];
      xy_part[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
  }
}

// ---------------------------------------------------------------- symmetric dense fp32
// Y = scale * A X + shift * X for a symmetric A reading only the 128-column strips on and
// above the diagonal: half the matrix bytes of the full product.  The upper part is cut
// into units of 64 rows x 128 columns, strip by strip (strip J: units s = 0 .. 2J+1, rows
// 64s .. 64s+63, columns 128J .. 128J+127; s >= 2J is the diagonal block), and the units are
// split evenly over a persistent grid of three blocks per CU.  Each wave of a block takes
// 16 rows of a unit (lane l: columns 2l, 2l+1 of every row):
//   column sums (K^T x for the rows' x: the block's mirror image below the diagonal) on the
//   VALU, accumulated over the block's consecutive units of one strip in registers and
//   summed over the four waves when the strip ends;
//   row sums (K x for the strip's x) on v_mfma_f32_16x16x4f32 from the wave's own 16 rows
//   staged in LDS (lane = row l % 16, four columns at 4 (l / 16)), written per unit.
// Diagonal units compute row sums over the whole 128 x 128 block and no column sums.
// dense_sym_reduce then adds, per 64 rows, the row sums of every strip to the right and the
// column sums of the rows above, in a fixed order (bit-reproducible).
constexpr int kSymCols = 128;
constexpr int kSymKs4 = 33;  // LDS row stride of a wave's 16 x 128 block in float4 (bank padding)
// persistent blocks: three per CU (a fourth block's 39.9 KB of LDS does not fit beside them;
// a grid sized for four left a tail of a quarter of the blocks: 143 against 126 us at 8 rhs)
constexpr int kSymGrid = 768;
constexpr int kSymNB = 8;    // rhs per pass

__host__ __device__ inline int64_t sym_strip_start(int64_t J) { return J * (J + 1); }
__host__ __device__ inline int64_t sym_units(int64_t nJ) { return sym_strip_start(nJ); }
__host__ __device__ inline int64_t sym_lo(int64_t w, int64_t U, int G) { return w * U / G; }
// block whose unit range holds unit u: the largest w with sym_lo(w) <= u
__host__ __device__ inline int sym_block_of(int64_t u, int64_t U, int G) { return (int)(((u + 1) * G - 1) / U); }
__host__ __device__ inline int sym_strip_of(int64_t u) {
  int64_t J = (int64_t)((sqrt(4.0 * (double)u + 1.0) - 1.0) * 0.5);
  while (sym_strip_start(J + 1) <= u) ++J;
  while (J > 0 && sym_strip_start(J) > u) --J;
  return (int)J;
}
inline int sym_grid(int64_t U) { return (int)(U < kSymGrid ? U : kSymGrid); }
// strips one block's range can touch: interior strips hold >= 4 units each
inline int sym_slots(int64_t U, int G) { return (int)(((U + G - 1) / G + 3) / 4 + 2); }

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// qr: units x NB x 64 row sums; cp: blocks x mslots x NB x 128 column sums
template <int NB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void dense_sym_kernel(
    int n, const float* __restrict__ A, int64_t lda, const float* __restrict__ X, int64_t ldx, int nbr,
    float* __restrict__ qr, float* __restrict__ cp, int mslots, const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ float4 ks4[4 * 16 * kSymKs4];
  __shared__ float4 xs4[8 * 4 * kSymNB];  // [column block][lane / 16][rhs]: x at 4 consecutive columns
  __shared__ float xis[4][16][kSymNB];    // per wave: x at its 16 rows, rhs fastest
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t U = sym_units(n / kSymCols);
  const int G = gridDim.x, w = blockIdx.x;
  const int64_t lo = sym_lo(w, U, G), hi = sym_lo(w + 1, U, G);
  if (lo >= hi) return;
  const int Jlo = sym_strip_of(lo);
  float* ksw = reinterpret_cast<float*>(ks4 + wv * 16 * kSymKs4);
  const int r16 = lane & 15, g = lane >> 4;
  float accz[2][NB];  // column sums at columns 2 lane, 2 lane + 1
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int b = 0; b < NB; ++b) accz[t][b] = 0.f;
  bool have = false;
  int curJ = -1;
  // the unit's rows as loaded and x at (row lane % 16, rhs lane / 16 and lane / 16 + 4); the
  // next unit's loads are issued as soon as the rows are in LDS
  float2 kr[16];
  float xv[2];
  // units are walked in order, so their strips follow from integer compares (no square root
  // per unit)
  auto load_unit = [&](int64_t u, int J) {
    const int64_t s = u - sym_strip_start(J);
    const float* base = A + (64 * s + 16 * wv) * lda + (int64_t)J * kSymCols + 2 * lane;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const f32x2_nt v = __builtin_nontemporal_load(reinterpret_cast<const f32x2_nt*>(base + i * lda));
      kr[i] = make_float2(v.x, v.y);
    }
    // x at the rows, loaded for every unit (diagonal ones do not use it) and masked where it is
    // used: a load behind a branch, or a select right after it, makes the wave wait for all of
    // its loads -- the next unit's rows included -- before the row-sum phase
    const int64_t r = 64 * s + 16 * wv + r16;
    const int b0 = g < nbr ? g : 0, b1 = g + 4 < nbr ? g + 4 : 0;  // clamped rhs
    xv[0] = X[(int64_t)b0 * ldx + r];
    xv[1] = X[(int64_t)b1 * ldx + r];
  };
  auto flush = [&](int slot) {
    __syncthreads();  // every wave is past its row sums: ks4 is free
    float* red = reinterpret_cast<float*>(ks4);
#pragma unroll
    for (int b = 0; b < NB; ++b)
      *reinterpret_cast<float2*>(red + (wv * NB + b) * kSymCols + 2 * lane) = make_float2(accz[0][b], accz[1][b]);
    __syncthreads();
    if (slot < mslots) {
      for (int e = threadIdx.x; e < NB * kSymCols; e += 256) {
        const int b = e / kSymCols, c = e % kSymCols;
        if (b < nbr)
          cp[((int64_t)(w * mslots + slot) * NB + b) * kSymCols + c] =
              ((red[(0 * NB + b) * kSymCols + c] + red[(1 * NB + b) * kSymCols + c]) +
               red[(2 * NB + b) * kSymCols + c]) + red[(3 * NB + b) * kSymCols + c];
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int b = 0; b < NB; ++b) accz[t][b] = 0.f;
  };
  int J = Jlo;
  load_unit(lo, J);
  for (int64_t u = lo; u < hi; ++u) {
    while (sym_strip_start(J + 1) <= u) ++J;
    const int64_t s = u - sym_strip_start(J);
    const bool diag = s >= 2 * (int64_t)J;
    if (have && (J != curJ || diag)) {
      flush(curJ - Jlo);
      have = false;
    }
    if (J != curJ) {
      __syncthreads();  // the previous strip's x is no longer read
      for (int e = threadIdx.x; e < 8 * 4 * kSymNB; e += 256) {
        const int r = e % kSymNB, gg = (e / kSymNB) % 4, cb = e / (4 * kSymNB);
        xs4[e] = r < nbr ? *reinterpret_cast<const float4*>(X + (int64_t)r * ldx + (int64_t)J * kSymCols + cb * 16 +
                                                             4 * gg)
                         : make_float4(0.f, 0.f, 0.f, 0.f);
      }
      __syncthreads();
      curJ = J;
    }
    if (!diag) {
      // column sums; the rows' x values go through LDS and come back wave-uniform
      if (g < NB) xis[wv][r16][g] = g < nbr ? xv[0] : 0.f;
      if (g + 4 < NB) xis[wv][r16][g + 4] = g + 4 < nbr ? xv[1] : 0.f;
      wave_lds_sync();
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        float xr[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) xr[b] = xis[wv][i][b];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          accz[0][b] = fmaf(kr[i].x, xr[b], accz[0][b]);
          accz[1][b] = fmaf(kr[i].y, xr[b], accz[1][b]);
        }
      }
      have = true;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) *reinterpret_cast<float2*>(ksw + i * 4 * kSymKs4 + 2 * lane) = kr[i];
    wave_lds_sync();
    if (u + 1 < hi) load_unit(u + 1, sym_strip_start(J + 1) <= u + 1 ? J + 1 : J);
    // row sums: D[row][rhs] += K[row][c] x[c][rhs] over the strip's 8 column blocks
    floatx4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
    const float4* ksr = reinterpret_cast<const float4*>(ksw) + r16 * kSymKs4 + g;
    // column j of D depends only on column j of B: lanes r16 >= 8 read another rhs's x (an
    // unconditional read) and fill columns that are never stored
    const float4* xsr = xs4 + g * kSymNB + (r16 & (kSymNB - 1));
#pragma unroll
    for (int cb = 0; cb < 8; cb += 2) {
      const float4 a0 = ksr[cb * 4];
      const float4 a1 = ksr[(cb + 1) * 4];
      const float4 b0 = xsr[cb * 4 * kSymNB];
      const float4 b1 = xsr[(cb + 1) * 4 * kSymNB];
      d0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0.x, b0.x, d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1.x, b1.x, d1, 0, 0, 0);
      d0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0.y, b0.y, d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1.y, b1.y, d1, 0, 0, 0);
      d0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0.z, b0.z, d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1.z, b1.z, d1, 0, 0, 0);
      d0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0.w, b0.w, d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1.w, b1.w, d1, 0, 0, 0);
    }
    // lane (g, r16) holds rows 4g .. 4g+3 of the wave's 16 for rhs r16
    if (r16 < nbr)
      *reinterpret_cast<float4*>(qr + (u * NB + r16) * 64 + 16 * wv + 4 * g) =
          make_float4(d0[0] + d1[0], d0[1] + d1[1], d0[2] + d1[2], d0[3] + d1[3]);
    wave_lds_sync();  // the rows are read before the next unit overwrites them
  }
  if (have) flush(curJ - Jlo);
}

// one block per 64 rows (sub-tile s of strip s / 2).  Wave q of eight sums the row sums of
// strips Ji + q, Ji + q + 8, ... (four strips' loads per batch, all rhs at once) and the column
// sums of every eighth block that covered strip Ji above the diagonal; then wave q takes rhs q:
// the eight waves' sums in order, the epilogue and the X . Y partials.
constexpr int kSymRedWaves = 8;
template <int NB>
__global__ __launch_bounds__(64 * kSymRedWaves) void dense_sym_reduce(int n, const float* __restrict__ X, int64_t ldx,
                                                        float* __restrict__ Y, int64_t ldy, int nbr, float scale,
                                                        float shift, const float* __restrict__ qr,
                                                        const float* __restrict__ cp, int mslots, int G,
                                                        const int* __restrict__ done, float* __restrict__ xy_part) {
  if (done && *done) return;
  __shared__ float red[kSymRedWaves][NB][64];
  const int s = blockIdx.x, nJ = n / kSymCols, Ji = s / 2;
  const int64_t U = sym_units(nJ);
  const int il = threadIdx.x & 63, q = threadIdx.x >> 6;
  float acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = 0.f;
  for (int J0 = Ji + q; J0 < nJ; J0 += 4 * kSymRedWaves) {
    float t[4][NB];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int J = J0 + kSymRedWaves * k < nJ ? J0 + kSymRedWaves * k : nJ - 1;  // clamped: unconditional load
      const float* src = qr + (sym_strip_start(J) + s) * NB * 64 + il;
#pragma unroll
      for (int b = 0; b < NB; ++b) t[k][b] = src[b * 64];
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const float v0 = J0 < nJ ? t[0][b] : 0.f, v1 = J0 + kSymRedWaves < nJ ? t[1][b] : 0.f;
      const float v2 = J0 + 2 * kSymRedWaves < nJ ? t[2][b] : 0.f;
      const float v3 = J0 + 3 * kSymRedWaves < nJ ? t[3][b] : 0.f;
      acc[b] += (v0 + v1) + (v2 + v3);
    }
  }
  // column sums of the blocks wf + q, wf + q + 4, ... that covered strip Ji above the diagonal
  if (Ji > 0) {
    const int wf = sym_block_of(sym_strip_start(Ji), U, G);
    const int wl = sym_block_of(sym_strip_start(Ji) + 2 * (int64_t)Ji - 1, U, G);
    for (int w = wf + q; w <= wl; w += kSymRedWaves) {
      const int slot = Ji - sym_strip_of(sym_lo(w, U, G));
      const float* src =
          cp + ((int64_t)(w * mslots + (slot < mslots ? slot : 0)) * NB) * kSymCols + 64 * (s % 2) + il;
      float t[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) t[b] = src[b * kSymCols];
#pragma unroll
      for (int b = 0; b < NB; ++b) acc[b] += slot < mslots ? t[b] : 0.f;
    }
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) red[q][b][il] = acc[b];
  __syncthreads();
  const int64_t row = 64 * (int64_t)s + il;
  for (int b = q; b < NB; b += kSymRedWaves) {
    if (b >= nbr) break;
    float sum = 0.f;
#pragma unroll
    for (int w = 0; w < kSymRedWaves; ++w) sum += red[w][b][il];
    const float x = X[(int64_t)b * ldx + row];
    float y = scale * sum;
    if (shift != 0.f) y = fmaf(shift, x, y);
    Y[(int64_t)b * ldy + row] = y;
    if (xy_part) {
      // two partials of 32 rows each, as the full product's epilogue writes them
      float v = x * y;
      v += dpp_f32<0xB1>(v);
      v += dpp_f32<0x4E>(v);
      v += dpp_f32<0x124>(v);
      v += dpp_f32<0x128>(v);
      // read the lanes before the branch: the last add must not move under the lane test
      const float h0 = readlane_f32(v, 0) + readlane_f32(v, 16);
      const float h1 = readlane_f32(v, 32) + readlane_f32(v, 48);
      if (il == 0) {
        xy_part[(int64_t)b * (n / 32) + 2 * s] = h0;
        xy_part[(int64_t)b * (n / 32) + 2 * s + 1] = h1;
      }
    }
  }
}

// scratch of the symmetric product, for up to kSymNB rhs per pass
size_t sym_scratch_bytes(int64_t n) {
  const int64_t U = sym_units(n / kSymCols);
  const int G = sym_grid(U);
  const size_t q = (size_t)U * kSymNB * 64 * sizeof(float), c = (size_t)G * sym_slots(U, G) * kSymNB * kSymCols * 4;
  return (q + 255) / 256 * 256 + (c + 255) / 256 * 256;
}

bool sym_eligible(const ck_op* op, const float* X, int64_t ldx, size_t avail) {
  auto al16 = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
  const int64_t n = op->n;
  return op->kind == CK_OP_DENSE_F32 && (op->flags & CK_OP_SYMMETRIC) && n >= kSymCols && n % kSymCols == 0 &&
         op->lda % 4 == 0 && ldx % 4 == 0 && al16(op->a) && al16(X) && avail >= sym_scratch_bytes(n);
}

// Y = op(X) through the symmetric kernels, in passes of up to eight rhs; scratch is the
// workspace past dense_sym_offset.  xy_part (optional, nrhs <= 8): dense_xy_nparts(n)
// partials of X . Y per rhs.
int dense_sym_apply(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, const int* done,
                    void* scratch, float* xy_part, hipStream_t s) {
  const int n = op->n;
  const int64_t U = sym_units(n / kSymCols);
  const int G = sym_grid(U), ms = sym_slots(U, G);
  const size_t qb = ((size_t)U * kSymNB * 64 * sizeof(float) + 255) / 256 * 256;
  float* qr = static_cast<float*>(scratch);
  float* cp = reinterpret_cast<float*>(static_cast<char*>(scratch) + qb);
  const float* A = static_cast<const float*>(op->a);
  for (int b0 = 0; b0 < nrhs; b0 += kSymNB) {
    const int nb = nrhs - b0 < kSymNB ? nrhs - b0 : kSymNB;
    const float* Xb = X + b0 * ldx;
    float* Yb = Y + b0 * ldy;
    float* xyb = xy_part && nrhs <= kSymNB ? xy_part : nullptr;
#define CK_SYM(NB_)                                                                                              \
  do {                                                                                                           \
    hipLaunchKernelGGL((dense_sym_kernel<NB_>), dim3(G), dim3(256), 0, s, n, A, op->lda, Xb, ldx, nb, qr, cp,  \
                       ms, done);                                                                                \
    hipLaunchKernelGGL((dense_sym_reduce<NB_>), dim3(n / 64), dim3(64 * kSymRedWaves), 0, s, n, Xb, ldx, Yb, ldy, nb, \
                       op->scale, op->shift, (const float*)qr, (const float*)cp, ms, G, done, xyb);              \
  } while (0)
    if (nb <= 1)
      CK_SYM(1);
    else if (nb <= 2)
      CK_SYM(2);
    else if (nb <= 4)
      CK_SYM(4);
    else
      CK_SYM(8);
#undef CK_SYM
  }
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int dense_dispatch(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, const int* done,
                   void* work, size_t work_bytes, hipStream_t s) {
  const size_t off = dense_sym_offset(op->n, nrhs);
  if (work && work_bytes > off && sym_eligible(op, X, ldx, work_bytes - off))
    return dense_sym_apply(op, nrhs, X, ldx, Y, ldy, done, static_cast<char*>(work) + off, nullptr, s);
  const int n = op->n;
  const bool bf16 = op->kind == CK_OP_DENSE_BF16;
  auto aligned16 = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
  // float4 / uint4 loads need 16-byte aligned rows and vector starts
  const bool vec = (n % 4 == 0) && (op->lda % 4 == 0) && (ldx % 4 == 0) && aligned16(op->a) && aligned16(X);
  if (bf16 && ((n % 8) || (op->lda % 8) || (ldx % 4) || !aligned16(op->a) || !aligned16(X)))
    return CK_ERR_UNSUPPORTED;
  // LDS-staged path for 3+ rhs when rows and vectors allow 16-byte loads
  const bool lds_ok = vec && (!bf16 || n % 8 == 0) && (ldx % 4 == 0) && nrhs >= 3;
  for (int b0 = 0; b0 < nrhs; b0 += 16) {
    const int nb = nrhs - b0 < 16 ? nrhs - b0 : 16;
    const float* Xb = X + b0 * ldx;
    float* Yb = Y + b0 * ldy;
    if (lds_ok && nb >= 3) {
#define CK_DENSE_LDS(NB_, ROWS_)                                                                                  \
  do {                                                                                                            \
    const int grid = (n + kDenseWaves * (ROWS_)-1) / (kDenseWaves * (ROWS_));                                     \
    if (bf16)                                                                                                     \
      hipLaunchKernelGGL((dense_lds_kernel<NB_, ROWS_, true>), dim3(grid), dim3(64 * kDenseWaves), 0, s, n,      \
                         op->a, op->lda, Xb, ldx, Yb, ldy, nb, op->scale, op->shift, done, (int64_t)0, (int64_t)0, \
                         (int64_t)0);                                                                             \
    else                                                                                                          \
      hipLaunchKernelGGL((dense_lds_kernel<NB_, ROWS_, false>), dim3(grid), dim3(64 * kDenseWaves), 0, s, n,     \
                         op->a, op->lda, Xb, ldx, Yb, ldy, nb, op->scale, op->shift, done, (int64_t)0, (int64_t)0, \
                         (int64_t)0);                                                                             \
  } while (0)
      if (nb <= 4)
        CK_DENSE_LDS(4, 4);
      else if (nb <= 8)
        CK_DENSE_LDS(8, 4);
      else
        CK_DENSE_LDS(16, 2);
#undef CK_DENSE_LDS
      continue;
    }
    // ROWS rows per wave: more rows reuse each loaded x chunk more often.
    constexpr int WAVES = 4;
#define CK_DENSE_LAUNCH(NB_, ROWS_)                                                                               \
  do {                                                                                                            \
    const int nwaves = (n + (ROWS_)-1) / (ROWS_);                                                                 \
    const int grid = (nwaves + WAVES - 1) / WAVES;                                                                \
    if (bf16)                                                                                                     \
      hipLaunchKernelGGL((dense_bf16_kernel<NB_, ROWS_>), dim3(grid), dim3(64 * WAVES), 0, s, n,                  \
                         (const uint16_t*)op->a, op->lda, Xb, ldx, Yb, ldy, nb, op->scale, op->shift, done);      \
    else                                                                                                          \
      hipLaunchKernelGGL((dense_f32_kernel<NB_, ROWS_>), dim3(grid), dim3(64 * WAVES), 0, s, n,                   \
                         (const float*)op->a, op->lda, Xb, ldx, Yb, ldy, nb, op->scale, op->shift, done, vec);    \
  } while (0)
    if (nb == 1)
      CK_DENSE_LAUNCH(1, 1);
    else if (nb == 2)
      CK_DENSE_LAUNCH(2, 2);
    else if (nb <= 4)
      CK_DENSE_LAUNCH(4, 4);
    else if (nb <= 8)
      CK_DENSE_LAUNCH(8, 4);
    else
      CK_DENSE_LAUNCH(16, 2);
#undef CK_DENSE_LAUNCH
  }
  CK_CHECK_LAUNCH();
  return CK_OK;
}

// ============================================================== CSR
__global__ __launch_bounds__(256) void csr_wave_per_row(int n, const int32_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ colind,
                                                        const float* __restrict__ val, const float* __restrict__ X,
                                                        int64_t ldx, float* __restrict__ Y, int64_t ldy, int nrhs,
                                                        float scale, float shift, const int* __restrict__ done) {
  if (done && *done) return;
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (row >= n) return;
  const int k0 = rowptr[row], k1 = rowptr[row + 1];
  for (int b = 0; b < nrhs; ++b) {
    const float* xb = X + b * ldx;
    float acc = 0.f;
    for (int k = k0 + lane; k < k1; k += 64) acc = fmaf(val[k], xb[colind[k]], acc);
    acc = wave_sum(acc);
    if (lane == 0) {
      float y = scale * acc;
      if (shift != 0.f) y = fmaf(shift, xb[row], y);
      Y[b * ldy + row] = y;
    }
  }
}

// Short rows: one lane per row.  A row of at most kCsrShort nonzeros is loaded into
// registers in one batch (column indices and values, then the x gathers of NB vectors),
// so each lane waits for memory twice per NB vectors instead of once per nonzero; the
// matrix is read once for all vectors.  Longer rows take a plain loop.  Every row is
// summed k ascending with one fma per nonzero, in both paths, so results do not depend
// on the path or on the vector count.
constexpr int kCsrShort = 8;

// XY: also xy_part[b * gridDim.x + block] = sum over the block's rows of X[b][row] Y[b][row] (a
// Krylov step's x . Ax from the matvec itself, instead of a separate pass over both vectors).
// The XY launch takes 1024 rows per block, so its partials are as few as a 1024-element tile
// pass would give: every block of the kernel that consumes them reads all of them.
template <int NB, bool XY = false>
__global__ __launch_bounds__(XY ? 1024 : 256) void csr_lane_per_row(int n, const int32_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ colind,
                                                        const float* __restrict__ val, const float* __restrict__ X,
                                                        int64_t ldx, float* __restrict__ Y, int64_t ldy, int nrhs,
                                                        float scale, float shift, const int* __restrict__ done,
                                                        float* __restrict__ xy_part) {
  if (done && *done) return;
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (!XY && row >= n) return;
  const int rc = row < n ? row : n - 1;  // XY: rows past n compute a copy of row n - 1 and store nothing
  const int k0 = rowptr[rc], k1 = rowptr[rc + 1];
  const int len = k1 - k0;
  if (len <= kCsrShort) {
    int c[kCsrShort];
    float v[kCsrShort];
#pragma unroll
    for (int i = 0; i < kCsrShort; ++i) {
      c[i] = i < len ? colind[k0 + i] : rc;
      v[i] = i < len ? val[k0 + i] : 0.f;
    }
    for (int b0 = 0; b0 < nrhs; b0 += NB) {
      float xv[NB][kCsrShort];
#pragma unroll
      for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int i = 0; i < kCsrShort; ++i)
          xv[j][i] = (i < len && b0 + j < nrhs) ? X[(int64_t)(b0 + j) * ldx + c[i]] : 0.f;
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const int b = b0 + j;
        if (b >= nrhs) break;
        float acc = 0.f;
#pragma unroll
        for (int i = 0; i < kCsrShort; ++i)
          if (i < len) acc = fmaf(v[i], xv[j][i], acc);
        float y = scale * acc;
        if (shift != 0.f) y = fmaf(shift, X[(int64_t)b * ldx + rc], y);
        if (!XY || row < n) Y[(int64_t)b * ldy + row] = y;
      }
    }
  } else {
    for (int b = 0; b < nrhs; ++b) {
      const float* xb = X + (int64_t)b * ldx;
      float acc = 0.f;
      for (int k = k0; k < k1; ++k) acc = fmaf(val[k], xb[colind[k]], acc);
      float y = scale * acc;
      if (shift != 0.f) y = fmaf(shift, xb[rc], y);
      if (!XY || row < n) Y[(int64_t)b * ldy + row] = y;
    }
  }
  if constexpr (XY) {
    __shared__ float red[16];
    for (int b = 0; b < nrhs; ++b) {
      const float xy = row < n ? X[(int64_t)b * ldx + row] * Y[(int64_t)b * ldy + row] : 0.f;
      const float sum = block_sum<1024>(xy, red);
      if (threadIdx.x == 0) xy_part[(int64_t)b * gridDim.x + blockIdx.x] = sum;
    }
  }
}

int csr_dispatch(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, const int* done,
                 hipStream_t s) {
  const int n = op->n;
  if (!op->rowptr || !op->colind || !op->values) return CK_ERR_BAD_ARG;
  const double avg = n > 0 ? (double)op->nnz / n : 0.0;
  if (avg > 32.0) {
    const int grid = (n + 3) / 4;
    hipLaunchKernelGGL(csr_wave_per_row, dim3(grid), dim3(256), 0, s, n, op->rowptr, op->colind, op->values, X, ldx,
                       Y, ldy, nrhs, op->scale, op->shift, done);
  } else {
    const dim3 grid((unsigned)((n + 255) / 256)), blk(256);
    if (nrhs == 1)
      hipLaunchKernelGGL(csr_lane_per_row<1>, grid, blk, 0, s, n, op->rowptr, op->colind, op->values, X, ldx, Y, ldy,
                         nrhs, op->scale, op->shift, done, nullptr);
    else if (nrhs == 2)
      hipLaunchKernelGGL(csr_lane_per_row<2>, grid, blk, 0, s, n, op->rowptr, op->colind, op->values, X, ldx, Y, ldy,
                         nrhs, op->scale, op->shift, done, nullptr);
    else
      hipLaunchKernelGGL(csr_lane_per_row<4>, grid, blk, 0, s, n, op->rowptr, op->colind, op->values, X, ldx, Y, ldy,
                         nrhs, op->scale, op->shift, done, nullptr);
  }
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // namespace

namespace ck {

int csr_apply_xy(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, const int* done,
                 float* xy_part, hipStream_t s) {
  const int n = op->n;
  if (op->kind != CK_OP_CSR_F32 || !op->rowptr || !op->colind || !op->values || !xy_part) return CK_ERR_BAD_ARG;
  if (n > 0 && (double)op->nnz / n > 32.0) return CK_ERR_UNSUPPORTED;  // the wave-per-row kernel has no epilogue
  const dim3 grid((unsigned)csr_xy_nparts(n)), blk(1024);
  if (nrhs == 1)
    hipLaunchKernelGGL((csr_lane_per_row<1, true>), grid, blk, 0, s, n, op->rowptr, op->colind, op->values, X, ldx, Y,
                       ldy, nrhs, op->scale, op->shift, done, xy_part);
  else if (nrhs == 2)
    hipLaunchKernelGGL((csr_lane_per_row<2, true>), grid, blk, 0, s, n, op->rowptr, op->colind, op->values, X, ldx, Y,
                       ldy, nrhs, op->scale, op->shift, done, xy_part);
  else
    hipLaunchKernelGGL((csr_lane_per_row<4, true>), grid, blk, 0, s, n, op->rowptr, op->colind, op->values, X, ldx, Y,
                       ldy, nrhs, op->scale, op->shift, done, xy_part);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // namespace ck

namespace {

// ============================================================== diagonal / identity
__global__ __launch_bounds__(256) void diag_kernel(int64_t n, int nrhs, const float* __restrict__ d,
                                                   const float* __restrict__ X, int64_t ldx, float* __restrict__ Y,
                                                   int64_t ldy, float scale, float shift,
                                                   const int* __restrict__ done) {
  if (done && *done) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (i >= n || b >= nrhs) return;
  const float x = X[b * ldx + i];
  const float s = d ? d[i] : 1.f;
  Y[b * ldy + i] = fmaf(scale * s, x, shift * x);
}

// ============================================================== Kronecker (MFMA)
constexpr int KP = 66;  // LDS row pitch (floats): conflict-free 16x4 operand reads

__device__ __forceinline__ floatx4 mfma16x16x4(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Load an r x c row-major global matrix (leading dim ld) into LDS [64][KP], zero padded.
__device__ __forceinline__ void load_pad64(float* __restrict__ dst, const float* __restrict__ src, int r, int c,
                                           int64_t ld) {
  for (int idx = threadIdx.x; idx < 64 * 64; idx += blockDim.x) {
    const int i = idx >> 6, j = idx & 63;
    dst[i * KP + j] = (i < r && j < c) ? src[(int64_t)i * ld + j] : 0.f;
  }
}

// Wave w computes rows [16w, 16w+16) x cols [0,64) of C = A * B^T where A, Bt are
// LDS [64][KP] (Bt holds B transposed, i.e. Bt[col][k]); k runs over kdim (<= 64).
__device__ __forceinline__ void wave_gemm_abt(const float* __restrict__ A, const float* __restrict__ Bt, int kdim,
                                              floatx4 acc[4]) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int ra = (wave * 16 + (lane & 15)) * KP;
  const int kk = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int k = 0; k < kdim; k += 4) {
    const float a = A[ra + k + kk];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const float b = Bt[(t * 16 + (lane & 15)) * KP + k + kk];
      acc[t] = mfma16x16x4(a, b, acc[t]);
    }
  }
}

// Pass A: for slab (b, i1): U = K2 * S * K3^T, S = X[b][i1] viewed as n2 x n3.
// LAST: U is the operator output (two-factor case): Y = scale*U + shift*S.
template <bool LAST>
__global__ __launch_bounds__(256) void kron_pass_a(int n1, int n2, int n3, const float* __restrict__ K2,
                                                   const float* __restrict__ K3, const float* __restrict__ X,
                                                   int64_t ldx, float* __restrict__ Out, int64_t ldo, float scale,
                                                   float shift, const int* __restrict__ done) {
  if (done && *done) return;
  extern __shared__ float lds[];
  float* S = lds;           // [64][KP] slab
  float* Kx = S + 64 * KP;  // K3 for the first product, then K2
  float* Tt = Kx + 64 * KP; // (S K3^T)^T
  const int i1 = blockIdx.x, b = blockIdx.y;
  const int64_t slab = (int64_t)b * ldx + (int64_t)i1 * n2 * n3;
  const int64_t oslab = (int64_t)b * ldo + (int64_t)i1 * n2 * n3;
  for (int idx = threadIdx.x; idx < 64 * 64; idx += blockDim.x) {
    const int i = idx >> 6, j = idx & 63;
    S[i * KP + j] = (i < n2 && j < n3) ? X[slab + i * n3 + j] : 0.f;
  }
  load_pad64(Kx, K3, n3, n3, n3);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  floatx4 acc[4];
  // T = S * K3^T  (T[r][c] = sum_k S[r][k] K3[c][k]); store transposed.
  wave_gemm_abt(S, Kx, n3, acc);
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int r = wave * 16 + 4 * (lane >> 4) + it;
      const int c = t * 16 + (lane & 15);
      Tt[c * KP + r] = acc[t][it];
    }
  __syncthreads();
  load_pad64(Kx, K2, n2, n2, n2);
  __syncthreads();
  // U = K2 * T  (U[r][c] = sum_k K2[r][k] T[k][c] = sum_k K2[r][k] Tt[c][k]).
  wave_gemm_abt(Kx, Tt, n2, acc);
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int r = wave * 16 + 4 * (lane >> 4) + it;
      const int c = t * 16 + (lane & 15);
      if (r < n2 && c < n3) {
        float u = acc[t][it];
        if (LAST) {
          u = scale * u;
          if (shift != 0.f) u = fmaf(shift, S[r * KP + c], u);
        }
        Out[oslab + r * n3 + c] = u;
      }
    }
}

// Pass B: Y[b][j1][c] = scale * sum_k K1[j1][k] U[b][k][c] + shift * X[b][j1][c]
// for a 64-column tile c in [64*blockIdx.x, +64).  Optionally accumulates the
// per-block partial sum of X*Y into partials[b * nparts + blockIdx.x].
__global__ __launch_bounds__(256) void kron_pass_b(int n1, int plane, const float* __restrict__ K1,
                                                   const float* __restrict__ U, int64_t ldu,
                                                   const float* __restrict__ X, int64_t ldx, float* __restrict__ Y,
                                                   int64_t ldy, float scale, float shift, float* __restrict__ partials,
                                                   int nparts, const int* __restrict__ done) {
  if (done && *done) return;
  extern __shared__ float lds[];
  float* Ka = lds;           // K1 [64][KP]
  float* Ut = Ka + 64 * KP;  // U tile transposed: Ut[c][k]
  __shared__ float red[4];
  const int b = blockIdx.y;
  const int c0 = blockIdx.x * 64;
  load_pad64(Ka, K1, n1, n1, n1);
  for (int idx = threadIdx.x; idx < 64 * 64; idx += blockDim.x) {
    const int k = idx >> 6, c = idx & 63;
    Ut[c * KP + k] = (k < n1 && c0 + c < plane) ? U[(int64_t)b * ldu + (int64_t)k * plane + c0 + c] : 0.f;
  }
  __syncthreads();
  floatx4 acc[4];
  wave_gemm_abt(Ka, Ut, n1, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float dot = 0.f;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int r = wave * 16 + 4 * (lane >> 4) + it;
      const int c = c0 + t * 16 + (lane & 15);
      if (r < n1 && c < plane) {
        const int64_t off = (int64_t)r * plane + c;
        const float x = X[(int64_t)b * ldx + off];
        float y = scale * acc[t][it];
        if (shift != 0.f) y = fmaf(shift, x, y);
        Y[(int64_t)b * ldy + off] = y;
        dot = fmaf(x, y, dot);
      }
    }
  if (partials) {
    const float s = block_sum<256>(dot, red);
    if (threadIdx.x == 0) partials[(int64_t)b * nparts + blockIdx.x] = s;
  }
}

}  // namespace

// ============================================================== internal API
namespace ck {

// Y = op(X) for a dense fp32 operator with 3..8 rhs on the LDS-staged kernel, writing the
// per-block partial sums of X . Y to xy_part (nrhs x dense_xy_nparts(n), column-major by
// rhs); returns CK_ERR_UNSUPPORTED when the shapes fall outside that kernel
int dense_apply_xy(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, const int* done,
                   float* xy_part, void* work, size_t work_bytes, hipStream_t s) {
  auto aligned16 = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
  const int n = op->n;
  const size_t off = dense_sym_offset(n, nrhs);
  if (work && work_bytes > off && nrhs >= 1 && nrhs <= 8 && sym_eligible(op, X, ldx, work_bytes - off))
    return dense_sym_apply(op, nrhs, X, ldx, Y, ldy, done, static_cast<char*>(work) + off, xy_part, s);
  if (op->kind != CK_OP_DENSE_F32 || nrhs < 3 || nrhs > 8 || n % 4 || op->lda % 4 || ldx % 4 || !aligned16(op->a) ||
      !aligned16(X))
    return CK_ERR_UNSUPPORTED;
  const int grid = dense_xy_nparts(n);
  if (nrhs <= 4)
    hipLaunchKernelGGL((dense_lds_kernel<4, 4, false>), dim3(grid), dim3(64 * kDenseWaves), 0, s, n, op->a, op->lda, X,
                       ldx, Y, ldy, nrhs, op->scale, op->shift, done, (int64_t)0, (int64_t)0, (int64_t)0, xy_part);
  else
    hipLaunchKernelGGL((dense_lds_kernel<8, 4, false>), dim3(grid), dim3(64 * kDenseWaves), 0, s, n, op->a, op->lda, X,
                       ldx, Y, ldy, nrhs, op->scale, op->shift, done, (int64_t)0, (int64_t)0, (int64_t)0, xy_part);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

size_t kron_workspace_bytes(const ck_op* op, int nrhs) {
  if (op->nfactors == 3) return (size_t)nrhs * op->n * sizeof(float);
  return 0;
}

int kron_dispatch(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, void* work,
                  size_t work_bytes, const int* done, float* partials, int nparts, hipStream_t s) {
  const int d = op->nfactors;
  if (d != 2 && d != 3) return CK_ERR_UNSUPPORTED;
  int64_t prod = 1;
  for (int i = 0; i < d; ++i) {
    if (op->fdim[i] < 1 || op->fdim[i] > 64 || !op->factor[i]) return CK_ERR_UNSUPPORTED;
    prod *= op->fdim[i];
  }
  if (prod != op->n) return CK_ERR_BAD_ARG;
  const size_t ldsA = 3 * 64 * KP * sizeof(float);
  const size_t ldsB = 2 * 64 * KP * sizeof(float);
  if (d == 2) {
    if (partials) return CK_ERR_UNSUPPORTED;
    hipLaunchKernelGGL((kron_pass_a<true>), dim3(1, nrhs), dim3(256), ldsA, s, 1, op->fdim[0], op->fdim[1],
                       op->factor[0], op->factor[1], X, ldx, Y, ldy, op->scale, op->shift, done);
    CK_CHECK_LAUNCH();
    return CK_OK;
  }
  if (work_bytes < kron_workspace_bytes(op, nrhs) || !work) return CK_ERR_WORKSPACE;
  if (kron64_eligible(op, X, ldx, Y, ldy))
    return kron64_apply_f32(op, nrhs, X, ldx, Y, ldy, work, work_bytes, done, partials, nparts, s);
  float* U = static_cast<float*>(work);
  const int n1 = op->fdim[0], n2 = op->fdim[1], n3 = op->fdim[2];
  const int plane = n2 * n3;
  hipLaunchKernelGGL((kron_pass_a<false>), dim3(n1, nrhs), dim3(256), ldsA, s, n1, n2, n3, op->factor[1],
                     op->factor[2], X, ldx, U, (int64_t)op->n, 1.f, 0.f, done);
  const int tiles = (plane + 63) / 64;
  if (partials && nparts != tiles) return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(kron_pass_b, dim3(tiles, nrhs), dim3(256), ldsB, s, n1, plane, op->factor[0], U,
                     (int64_t)op->n, X, ldx, Y, ldy, op->scale, op->shift, partials, nparts, done);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int kron_nparts(const ck_op* op) {
  if (op->nfactors != 3) return 0;
  return (op->fdim[1] * op->fdim[2] + 63) / 64;
}

int op_apply_internal(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, void* work,
                      size_t work_bytes, const int* done, hipStream_t s) {
  if (!op || !X || !Y || nrhs < 1 || op->n < 1 || ldx < op->n || ldy < op->n) return CK_ERR_BAD_ARG;
  switch (op->kind) {
    case CK_OP_IDENTITY:
    case CK_OP_DIAG_F32: {
      const float* d = op->kind == CK_OP_DIAG_F32 ? (const float*)op->a : nullptr;
      if (op->kind == CK_OP_DIAG_F32 && !d) return CK_ERR_BAD_ARG;
      dim3 grid((unsigned)cdiv(op->n, 256), nrhs);
      hipLaunchKernelGGL(diag_kernel, grid, dim3(256), 0, s, (int64_t)op->n, nrhs, d, X, ldx, Y, ldy, op->scale,
                         op->shift, done);
      CK_CHECK_LAUNCH();
      return CK_OK;
    }
    case CK_OP_DENSE_F32:
    case CK_OP_DENSE_BF16:
      if (!op->a || op->lda < op->n) return CK_ERR_BAD_ARG;
      return dense_dispatch(op, nrhs, X, ldx, Y, ldy, done, work, work_bytes, s);
    case CK_OP_DENSE_BATCHED_F32: {
      // vector b multiplies matrix b (a + b * nnz elements): independent dense systems of one order
      auto aligned16 = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
      if (!op->a || op->lda < op->n || op->nnz < (int64_t)op->lda * (op->n - 1) + op->n) return CK_ERR_BAD_ARG;
      if (op->n % 4 || op->lda % 4 || op->nnz % 4 || ldx % 4 || ldy % 4 || !aligned16(op->a) || !aligned16(X))
        return CK_ERR_UNSUPPORTED;
      const int grid = (op->n + kDenseWaves * 4 - 1) / (kDenseWaves * 4);
      hipLaunchKernelGGL((dense_lds_kernel<1, 4, false>), dim3(grid, nrhs), dim3(64 * kDenseWaves), 0, s, op->n,
                         op->a, op->lda, X, ldx, Y, ldy, 1, op->scale, op->shift, done, op->nnz, ldx, ldy);
      CK_CHECK_LAUNCH();
      return CK_OK;
    }
    case CK_OP_CSR_F32:
      return csr_dispatch(op, nrhs, X, ldx, Y, ldy, done, s);
    case CK_OP_KRON_F32:
      return kron_dispatch(op, nrhs, X, ldx, Y, ldy, work, work_bytes, done, nullptr, 0, s);
    case CK_OP_WOODBURY_INV:
      return woodbury_apply(op, nrhs, X, ldx, Y, ldy, work, work_bytes, done, s);
    case CK_OP_RBF_F32:
      return rbf_apply(op, nrhs, X, ldx, Y, ldy, done, s);
    case CK_OP_KRON_EIG_F32:
      if (!kron64_eligible(op, X, ldx, Y, ldy)) return CK_ERR_UNSUPPORTED;
      return kron64_eig_apply_f32(op, nrhs, X, ldx, Y, ldy, work, work_bytes, done, s);
    default:
      return CK_ERR_UNSUPPORTED;
  }
}

size_t op_workspace_internal(const ck_op* op, int nrhs) {
  if (!op) return 0;
  switch (op->kind) {
    case CK_OP_KRON_F32:
      return kron_workspace_bytes(op, nrhs);
    case CK_OP_WOODBURY_INV:
      return woodbury_workspace_bytes(op, nrhs);
    case CK_OP_KRON_EIG_F32:
      return 2 * (size_t)nrhs * op->n * sizeof(float);
    case CK_OP_DENSE_F32:
      // a symmetric operator's product reads half the matrix through scratch placed past the
      // regions CG's fused steps use (dense_sym_offset)
      if ((op->flags & CK_OP_SYMMETRIC) && op->n >= kSymCols && op->n % kSymCols == 0)
        return dense_sym_offset(op->n, nrhs) + sym_scratch_bytes(op->n);
      // room for the p.Ap partials a CG step takes from the matvec epilogue (dense_apply_xy)
      return (nrhs >= 3 && nrhs <= 8) ? dense_xy_bytes(op->n, nrhs) : 0;
    case CK_OP_CSR_F32:
      // the x.Ax partials of csr_apply_xy (CG's p.Ap, Lanczos's v.Av)
      return csr_xy_bytes(op->n, nrhs);
    default:
      return 0;
  }
}

}  // namespace ck

// ============================================================== transpose / dot
namespace {

__global__ __launch_bounds__(256) void transpose_kernel(int64_t n, int nrhs, const float* __restrict__ src,
                                                        int64_t lds_, float* __restrict__ dst, int64_t ldd,
                                                        int to_blocks) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (to_blocks) {
    for (int b = 0; b < nrhs; ++b) dst[b * ldd + i] = src[i * lds_ + b];
  } else {
    for (int b = 0; b < nrhs; ++b) dst[i * ldd + b] = src[b * lds_ + i];
  }
}

__global__ __launch_bounds__(256) void dot_partials_kernel(int64_t n, const float* __restrict__ X, int64_t ldx,
                                                           const float* __restrict__ Y, int64_t ldy,
                                                           double* __restrict__ part, int nparts) {
  __shared__ double red[4];
  const int b = blockIdx.y;
  const int64_t per = cdiv(n, nparts);
  const int64_t i0 = blockIdx.x * per, i1 = i0 + per < n ? i0 + per : n;
  double acc = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x)
    acc += (double)X[b * ldx + i] * (double)Y[b * ldy + i];
  acc = block_sum<256>(acc, red);
  if (threadIdx.x == 0) part[b * nparts + blockIdx.x] = acc;
}

__global__ __launch_bounds__(64) void dot_final_kernel(int nparts, const double* __restrict__ part,
                                                       double* __restrict__ out) {
  const int b = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 64) acc += part[b * nparts + i];
  acc = wave_sum(acc);
  if (threadIdx.x == 0) out[b] = acc;
}

constexpr int kDotParts = 256;

}  // namespace

extern "C" {

size_t ck_op_workspace_bytes(const ck_op* op, int32_t nrhs) { return ck::op_workspace_internal(op, nrhs); }

int ck_op_apply(const ck_op* op, int32_t nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, void* work,
                size_t work_bytes, void* stream) {
  return ck::op_apply_internal(op, nrhs, X, ldx, Y, ldy, work, work_bytes, nullptr, (hipStream_t)stream);
}

int ck_transpose_f32(int64_t n, int32_t nrhs, const float* src, int64_t lds_, float* dst, int64_t ldd,
                     int32_t to_blocks, void* stream) {
  if (n < 1 || nrhs < 1 || !src || !dst) return CK_ERR_BAD_ARG;
  if (to_blocks ? (lds_ < nrhs || ldd < n) : (lds_ < n || ldd < nrhs)) return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(transpose_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, n, nrhs, src,
                     lds_, dst, ldd, to_blocks);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

size_t ck_dot_workspace_bytes(int64_t n, int32_t nrhs) { return (size_t)nrhs * kDotParts * sizeof(double); }

int ck_dot_f32(int64_t n, int32_t nrhs, const float* X, int64_t ldx, const float* Y, int64_t ldy, double* out,
               void* work, size_t work_bytes, void* stream) {
  if (n < 1 || nrhs < 1 || !X || !Y || !out || ldx < n || ldy < n) return CK_ERR_BAD_ARG;
  if (!work || work_bytes < ck_dot_workspace_bytes(n, nrhs)) return CK_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  double* part = static_cast<double*>(work);
  hipLaunchKernelGGL(dot_partials_kernel, dim3(kDotParts, nrhs), dim3(256), 0, s, n, X, ldx, Y, ldy, part, kDotParts);
  hipLaunchKernelGGL(dot_final_kernel, dim3(nrhs), dim3(64), 0, s, kDotParts, part, out);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // extern "C"
