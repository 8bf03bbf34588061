// Mixed precision: iterative refinement with FP64 residuals around FP32 (or
// BF16-matrix) inner solves, and bf16 storage conversion.
//
//   ck_ir_residual_f64: r = b - Ax in fp64, its fp32 copy for the inner solve,
//                       and per-column ||r||^2, ||b||^2 partials -> fp64 norms
//   ck_ir_update_f64:   x += d (fp32 correction accumulated in fp64)
//   ck_ir_inner_start / ck_ir_update_scaled_f64: the inner CG's column normalisation
//                       and un-normalisation folded into the refinement passes
#include "common.hpp"

using namespace ck;

namespace {

constexpr int kIrParts = 256;

__global__ __launch_bounds__(256) void ir_residual_kernel(int64_t n, const double* __restrict__ b, int64_t ldb,
                                                          const double* __restrict__ Ax, int64_t ldax,
                                                          double* __restrict__ r, int64_t ldr,
                                                          float* __restrict__ r32, int64_t ldr32,
                                                          double* __restrict__ part) {
  __shared__ double red[4];
  const int col = blockIdx.y, blk = blockIdx.x;
  const int64_t per = cdiv(n, kIrParts);
  const int64_t i0 = blk * per, i1 = i0 + per < n ? i0 + per : n;
  double rr = 0.0, bb = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
    const double bv = b[col * ldb + i];
    const double rv = bv - (Ax ? Ax[col * ldax + i] : 0.0);
    if (r) r[col * ldr + i] = rv;
    if (r32) r32[col * ldr32 + i] = (float)rv;
    rr += rv * rv;
    bb += bv * bv;
  }
  rr = block_sum<256>(rr, red);
  bb = block_sum<256>(bb, red);
  if (threadIdx.x == 0) {
    part[(2 * col) * kIrParts + blk] = rr;
    part[(2 * col + 1) * kIrParts + blk] = bb;
  }
}

__global__ __launch_bounds__(64) void ir_norms_kernel(const double* __restrict__ part, double* __restrict__ out) {
  const int col = blockIdx.x, which = blockIdx.y;
  double acc = 0.0;
  for (int i = threadIdx.x; i < kIrParts; i += 64) acc += part[(2 * col + which) * kIrParts + i];
  acc = wave_sum(acc);
  if (threadIdx.x == 0) out[2 * col + which] = sqrt(acc);
}

__global__ __launch_bounds__(256) void ir_update_kernel(int64_t n, double* __restrict__ x, int64_t ldx,
                                                        const float* __restrict__ d, int64_t ldd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int col = blockIdx.y;
  if (i >= n) return;
  x[col * ldx + i] += (double)d[col * ldd + i];
}

// Start of a refinement solve from the caller's right-hand sides b (element (i, c) at
// b[i * rs + c * cs], fp32): b64 = b, r32 = b, x64 = 0 in the solver's column blocks, and
// the ||b||^2 partials of ir_residual_kernel for both norms (r = b at x = 0).  One pass
// over b instead of a transposing copy, a fill and a residual pass.  Up to kLoadCols
// columns per pass; a thread owns one row.
constexpr int kLoadCols = 16;

// FULL: the chunk has all kLoadCols columns and ZX: x64 is zeroed; both known on the host,
// so the per-column and per-store tests compile away (this kernel runs once per launch on
// a cold instruction cache, where every branch costs).
template <bool V4, bool FULL, bool ZX>
__global__ __launch_bounds__(256) void ir_load_rhs_kernel(int64_t n, int nrhs, const float* __restrict__ b,
                                                          int64_t rs, int64_t cs, double* __restrict__ b64,
                                                          float* __restrict__ r32, double* __restrict__ x64,
                                                          int64_t ld, int c0, double* __restrict__ part) {
  __shared__ double red[4];
  const int blk = blockIdx.x;
  const int64_t per = cdiv(n, kIrParts);
  const int64_t i0 = blk * per, i1 = i0 + per < n ? i0 + per : n;
  const int nc = FULL ? kLoadCols : (nrhs - c0 < kLoadCols ? nrhs - c0 : kLoadCols);
  double acc[kLoadCols];
#pragma unroll
  for (int j = 0; j < kLoadCols; ++j) acc[j] = 0.0;
  // rows i, i + 256, ... of this thread, kRowsInFlight at a time: every row's loads are
  // issued (index clamped into the block's range) before the first store
  constexpr int kRowsInFlight = 4;
  for (int64_t ib = i0 + threadIdx.x; ib < i1; ib += 256 * kRowsInFlight) {
    float row[kRowsInFlight][kLoadCols];
#pragma unroll
    for (int u = 0; u < kRowsInFlight; ++u) {
      const int64_t i = ib + 256 * u < i1 ? ib + 256 * u : i1 - 1;
      if (V4) {
        // a row's columns are contiguous and 16-byte aligned (cs == 1, rs and c0 multiples of 4)
#pragma unroll
        for (int j = 0; j < kLoadCols; j += 4)
          if (j < nc) {
            const float4 q = *reinterpret_cast<const float4*>(b + i * rs + c0 + j);
            row[u][j] = q.x;
            row[u][j + 1] = q.y;
            row[u][j + 2] = q.z;
            row[u][j + 3] = q.w;
          }
      } else {
#pragma unroll
        for (int j = 0; j < kLoadCols; ++j)
          if (j < nc) row[u][j] = b[i * rs + (int64_t)(c0 + j) * cs];
      }
    }
#pragma unroll
    for (int u = 0; u < kRowsInFlight; ++u) {
      const int64_t i = ib + 256 * u;
      if (i >= i1) break;
#pragma unroll
      for (int j = 0; j < kLoadCols; ++j) {
        if (j < nc) {
          const int c = c0 + j;
          const float v = row[u][j];
          const double dv = (double)v;
          b64[c * ld + i] = dv;
          r32[c * ld + i] = v;
          if (ZX) x64[c * ld + i] = 0.0;
          acc[j] += dv * dv;
        }
      }
    }
  }
  if (!part) return;  // uniform: no norms requested
#pragma unroll
  for (int j = 0; j < kLoadCols; ++j) {
    if (j < nc) {
      const double t = block_sum<256>(acc[j], red);
      if (threadIdx.x == 0) {
        part[(2 * (c0 + j)) * kIrParts + blk] = t;
        part[(2 * (c0 + j) + 1) * kIrParts + blk] = t;
      }
    }
  }
}

// Start of an inner correction solve on the refinement residual: r = r32 / safe(||r||),
// x = 0, mult = ||r|| (fp32) -- the column normalisation of the inner CG done with the
// FP64 norm the residual pass already produced.
__global__ __launch_bounds__(256) void ir_inner_start_kernel(int64_t n, const float* __restrict__ r32, int64_t ld32,
                                                             const double* __restrict__ norms, float* __restrict__ r,
                                                             float* __restrict__ x, int64_t ld,
                                                             float* __restrict__ mult) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int col = blockIdx.y;
  const float m = (float)norms[2 * col];
  if (i == 0) mult[col] = m;
  if (i >= n) return;
  const float ms = fabsf(m) < 1e-40f ? 1e-40f : m;
  r[col * ld + i] = r32[col * ld32 + i] / ms;
  x[col * ld + i] = 0.f;
}

// x += mult * d: the inner solve's un-normalisation (fp32 product) folded into the FP64 update
__global__ __launch_bounds__(256) void ir_update_scaled_kernel(int64_t n, double* __restrict__ x, int64_t ldx,
                                                               const float* __restrict__ d, int64_t ldd,
                                                               const float* __restrict__ mult) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int col = blockIdx.y;
  if (i >= n) return;
  x[col * ldx + i] += (double)(d[col * ldd + i] * mult[col]);
}

__global__ __launch_bounds__(256) void to_bf16_kernel(int64_t count, const float* __restrict__ src,
                                                      uint16_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  // round to nearest even (finite inputs)
  const uint32_t u = __float_as_uint(src[i]);
  const uint32_t rounded = u + 0x7FFFu + ((u >> 16) & 1u);
  dst[i] = (uint16_t)(rounded >> 16);
}

// float4 copy with four independent loads in flight per thread: the bandwidth
// probe the roofline peak and the PMC byte counters are calibrated with
constexpr int kCopyUnroll = 4;
typedef float f32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void stream_copy_kernel(int64_t n4, const f32x4* __restrict__ src,
                                                          f32x4* __restrict__ dst) {
  const int64_t base = (int64_t)blockIdx.x * (256 * kCopyUnroll) + threadIdx.x;
  f32x4 v[kCopyUnroll];
#pragma unroll
  for (int u = 0; u < kCopyUnroll; ++u) {
    const int64_t i = base + 256 * u;
    if (i < n4) v[u] = __builtin_nontemporal_load(src + i);
  }
#pragma unroll
  for (int u = 0; u < kCopyUnroll; ++u) {
    const int64_t i = base + 256 * u;
    if (i < n4) __builtin_nontemporal_store(v[u], dst + i);
  }
}

}  // namespace

extern "C" {

int ck_stream_copy_f32(int64_t n, const float* src, float* dst, void* stream) {
  if (n < 4 || (n % 4) || !src || !dst || ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15))
    return CK_ERR_BAD_ARG;
  const int64_t n4 = n / 4;
  const int64_t blocks = cdiv(n4, 256 * kCopyUnroll);
  if (blocks > 0x7FFFFFFF) return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(stream_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, n4,
                     reinterpret_cast<const f32x4*>(src), reinterpret_cast<f32x4*>(dst));
  CK_CHECK_LAUNCH();
  return CK_OK;
}

size_t ck_ir_workspace_bytes(int32_t nrhs) { return sizeof(double) * 2 * (size_t)nrhs * kIrParts; }

int ck_ir_residual_f64(int64_t n, int32_t nrhs, const double* b, int64_t ldb, const double* Ax, int64_t ldax,
                       double* r, int64_t ldr, float* r32, int64_t ldr32, double* norms, void* work,
                       size_t work_bytes, void* stream) {
  if (n < 1 || nrhs < 1 || !b || !norms || ldb < n || (Ax && ldax < n) || (r && ldr < n) || (r32 && ldr32 < n))
    return CK_ERR_BAD_ARG;
  if (!work || work_bytes < ck_ir_workspace_bytes(nrhs)) return CK_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  double* part = static_cast<double*>(work);
  hipLaunchKernelGGL(ir_residual_kernel, dim3(kIrParts, nrhs), dim3(256), 0, s, n, b, ldb, Ax, ldax, r, ldr, r32,
                     ldr32, part);
  hipLaunchKernelGGL(ir_norms_kernel, dim3(nrhs, 2), dim3(64), 0, s, (const double*)part, norms);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_ir_update_f64(int64_t n, int32_t nrhs, double* x, int64_t ldx, const float* d, int64_t ldd, void* stream) {
  if (n < 1 || nrhs < 1 || !x || !d || ldx < n || ldd < n) return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(ir_update_kernel, dim3((unsigned)cdiv(n, 256), nrhs), dim3(256), 0, (hipStream_t)stream, n, x,
                     ldx, d, ldd);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_ir_load_rhs(int64_t n, int32_t nrhs, const float* b, int64_t rs, int64_t cs, double* b64, float* r32,
                   double* x64, int64_t ld, double* norms, void* work, size_t work_bytes, void* stream) {
  if (n < 1 || nrhs < 1 || !b || !b64 || !r32 || ld < n) return CK_ERR_BAD_ARG;
  if (norms && (!work || work_bytes < ck_ir_workspace_bytes(nrhs))) return CK_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  double* part = norms ? static_cast<double*>(work) : nullptr;
  const bool v4 = cs == 1 && rs % 4 == 0 && nrhs % 4 == 0 && (reinterpret_cast<uintptr_t>(b) & 15) == 0;
  for (int c0 = 0; c0 < nrhs; c0 += kLoadCols) {
    const bool full = nrhs - c0 >= kLoadCols;
    auto launch = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(kIrParts), dim3(256), 0, s, n, nrhs, b, rs, cs, b64, r32, x64, ld, c0, part);
    };
    if (v4 && full)
      x64 ? launch(ir_load_rhs_kernel<true, true, true>) : launch(ir_load_rhs_kernel<true, true, false>);
    else if (v4)
      x64 ? launch(ir_load_rhs_kernel<true, false, true>) : launch(ir_load_rhs_kernel<true, false, false>);
    else
      x64 ? launch(ir_load_rhs_kernel<false, false, true>) : launch(ir_load_rhs_kernel<false, false, false>);
  }
  if (norms) hipLaunchKernelGGL(ir_norms_kernel, dim3(nrhs, 2), dim3(64), 0, s, (const double*)part, norms);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_ir_inner_start(int64_t n, int32_t nrhs, const float* r32, int64_t ld32, const double* norms, float* r,
                      float* x, int64_t ld, float* mult, void* stream) {
  if (n < 1 || nrhs < 1 || !r32 || !norms || !r || !x || !mult || ld32 < n || ld < n) return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(ir_inner_start_kernel, dim3((unsigned)cdiv(n, 256), nrhs), dim3(256), 0, (hipStream_t)stream, n,
                     r32, ld32, norms, r, x, ld, mult);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_ir_update_scaled_f64(int64_t n, int32_t nrhs, double* x, int64_t ldx, const float* d, int64_t ldd,
                            const float* mult, void* stream) {
  if (n < 1 || nrhs < 1 || !x || !d || !mult || ldx < n || ldd < n) return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(ir_update_scaled_kernel, dim3((unsigned)cdiv(n, 256), nrhs), dim3(256), 0, (hipStream_t)stream,
                     n, x, ldx, d, ldd, mult);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_f32_to_bf16(int64_t count, const float* src, uint16_t* dst, void* stream) {
  if (count < 1 || !src || !dst) return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)cdiv(count, 256)), dim3(256), 0, (hipStream_t)stream, count, src,
                     dst);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // extern "C"
