// Low-rank (Woodbury) inverse application and the on-the-fly RBF kernel operator.
//
// Woodbury: (s I + F F^T)^{-1} x = (x - F (s I + F^T F)^{-1} F^T x) / s with the
// small core inverse precomputed (rank x rank, fp64).  All arithmetic after the
// fp32 loads is fp64: for x in span(F) the subtraction cancels to a relative size
// s / (s + lambda), so fp32 arithmetic would leave errors of u * lambda / s there
// (0.06 at lambda / s = 1e6), enough to stall PCG.  F^T x is a tall-skinny
// reduction (one partial per 1024-row tile, added in index order); the core
// product and the rank-k update are fused into one pass over F.
//
// RBF: S_ij = exp(-|p_i - p_j|^2 / (2 l^2)) is never stored; each block stages
// 256 points and their x values in LDS and every lane accumulates one row
// (compute bound: one exp per pair).
#include "common.hpp"
#include "ops_internal.hpp"

using namespace ck;

namespace {

constexpr int kWTile = 1024;
constexpr int kMaxRank = 128;

__global__ __launch_bounds__(256) void wb_ftx_partials(int64_t n, int k, const float* __restrict__ F, int64_t lda,
                                                       const float* __restrict__ X, int64_t ldx,
                                                       double* __restrict__ part, int P,
                                                       const int* __restrict__ done) {
  if (done && *done) return;
  const int b = blockIdx.y, blk = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t i0 = (int64_t)blk * kWTile;
  float xv[kWTile / 64];
#pragma unroll
  for (int q = 0; q < kWTile / 64; ++q) {
    const int64_t i = i0 + lane + 64 * q;
    xv[q] = i < n ? X[b * ldx + i] : 0.f;
  }
  for (int c = wave; c < k; c += 4) {
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < kWTile / 64; ++q) {
      const int64_t i = i0 + lane + 64 * q;
      if (i < n) acc = fma((double)F[i * lda + c], (double)xv[q], acc);
    }
    acc = wave_sum(acc);
    if (lane == 0) part[((int64_t)b * k + c) * P + blk] = acc;
  }
}

__global__ __launch_bounds__(256) void wb_reduce(int k, const double* __restrict__ part, int P, double* __restrict__ t,
                                                 const int* __restrict__ done) {
  if (done && *done) return;
  const int b = blockIdx.y;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= k) return;
  double acc = 0.0;
  for (int p = threadIdx.x & 63; p < P; p += 64) acc += part[((int64_t)b * k + c) * P + p];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) t[b * k + c] = acc;
}

__global__ __launch_bounds__(256) void wb_apply(int64_t n, int k, const float* __restrict__ F, int64_t lda,
                                                const double* __restrict__ Cinv, const double* __restrict__ t,
                                                const float* __restrict__ X, int64_t ldx, float* __restrict__ Y,
                                                int64_t ldy, double scale, double shift, const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ double ts[kMaxRank], us[kMaxRank];
  const int b = blockIdx.y;
  for (int c = threadIdx.x; c < k; c += 256) ts[c] = t[b * k + c];
  __syncthreads();
  for (int c = threadIdx.x; c < k; c += 256) {
    double acc = 0.0;
    for (int j = 0; j < k; ++j) acc = fma(Cinv[c * k + j], ts[j], acc);
    us[c] = acc;
  }
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  const float* Fi = F + i * lda;
  for (int c = 0; c < k; ++c) acc = fma((double)Fi[c], us[c], acc);
  const double x = (double)X[b * ldx + i];
  Y[b * ldy + i] = (float)(scale * (x - acc) + shift * x);
}

constexpr int kRbfTile = 256;
constexpr int kRbfMaxDim = 8;
constexpr int kRbfNB = 8;

__global__ __launch_bounds__(256) void rbf_kernel(int n, int dim, const float* __restrict__ P, int64_t ldp,
                                                  float inv2l2, const float* __restrict__ X, int64_t ldx,
                                                  float* __restrict__ Y, int64_t ldy, int nrhs, float scale,
                                                  float shift, const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ float pj[kRbfTile][kRbfMaxDim];
  __shared__ float xj[kRbfNB][kRbfTile];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int b0 = blockIdx.y * kRbfNB;
  const int nb = nrhs - b0 < kRbfNB ? nrhs - b0 : kRbfNB;
  float pi[kRbfMaxDim];
#pragma unroll
  for (int d = 0; d < kRbfMaxDim; ++d) pi[d] = (i < n && d < dim) ? P[(int64_t)i * ldp + d] : 0.f;
  float acc[kRbfNB];
#pragma unroll
  for (int b = 0; b < kRbfNB; ++b) acc[b] = 0.f;
  for (int j0 = 0; j0 < n; j0 += kRbfTile) {
    __syncthreads();
    {
      const int j = j0 + threadIdx.x;
#pragma unroll
      for (int d = 0; d < kRbfMaxDim; ++d) pj[threadIdx.x][d] = (j < n && d < dim) ? P[(int64_t)j * ldp + d] : 0.f;
#pragma unroll
      for (int b = 0; b < kRbfNB; ++b) xj[b][threadIdx.x] = (j < n && b < nb) ? X[(int64_t)(b0 + b) * ldx + j] : 0.f;
    }
    __syncthreads();
    const int jn = n - j0 < kRbfTile ? n - j0 : kRbfTile;
    for (int jj = 0; jj < jn; ++jj) {
      float d2 = 0.f;
#pragma unroll
      for (int d = 0; d < kRbfMaxDim; ++d) {
        const float df = pi[d] - pj[jj][d];
        d2 = fmaf(df, df, d2);
      }
      const float kv = expf(-d2 * inv2l2);
#pragma unroll
      for (int b = 0; b < kRbfNB; ++b) acc[b] = fmaf(kv, xj[b][jj], acc[b]);
    }
  }
  if (i >= n) return;
  for (int b = 0; b < nb; ++b) {
    const float x = X[(int64_t)(b0 + b) * ldx + i];
    Y[(int64_t)(b0 + b) * ldy + i] = fmaf(shift, x, scale * acc[b]);
  }
}

}  // namespace

namespace ck {

size_t woodbury_workspace_bytes(const ck_op* op, int nrhs) {
  const size_t P = cdiv(op->n, kWTile);
  return sizeof(double) * ((size_t)nrhs * op->rank * P + (size_t)nrhs * op->rank);
}

int woodbury_apply(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, void* work,
                   size_t work_bytes, const int* done, hipStream_t s) {
  const int k = op->rank;
  if (k < 1 || k > kMaxRank || !op->a || !op->aux || op->lda < k) return CK_ERR_BAD_ARG;
  if (!work || work_bytes < woodbury_workspace_bytes(op, nrhs)) return CK_ERR_WORKSPACE;
  const int64_t n = op->n;
  const int P = (int)cdiv(n, kWTile);
  double* part = static_cast<double*>(work);
  double* t = part + (size_t)nrhs * k * P;
  const float* F = static_cast<const float*>(op->a);
  hipLaunchKernelGGL(wb_ftx_partials, dim3(P, nrhs), dim3(256), 0, s, n, k, F, op->lda, X, ldx, part, P, done);
  hipLaunchKernelGGL(wb_reduce, dim3((unsigned)cdiv(k, 4), nrhs), dim3(256), 0, s, k, (const double*)part, P, t,
                     done);
  hipLaunchKernelGGL(wb_apply, dim3((unsigned)cdiv(n, 256), nrhs), dim3(256), 0, s, n, k, F, op->lda,
                     reinterpret_cast<const double*>(op->aux), (const double*)t, X, ldx, Y, ldy, (double)op->scale,
                     (double)op->shift, done);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int rbf_apply(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, const int* done,
              hipStream_t s) {
  if (!op->a || op->dim < 1 || op->dim > kRbfMaxDim || op->lda < op->dim || !(op->aux_f > 0.f)) return CK_ERR_BAD_ARG;
  const float inv2l2 = 1.f / (2.f * op->aux_f * op->aux_f);
  hipLaunchKernelGGL(rbf_kernel, dim3((unsigned)cdiv(op->n, 256), (unsigned)cdiv(nrhs, kRbfNB)), dim3(256), 0, s,
                     op->n, op->dim, static_cast<const float*>(op->a), op->lda, inv2l2, X, ldx, Y, ldy, nrhs,
                     op->scale, op->shift, done);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // namespace ck
