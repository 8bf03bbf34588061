// Low-rank (Woodbury) inverse application and the on-the-fly RBF kernel operator.
//
// Woodbury: (s I + F F^T)^{-1} x = (x - F (s I + F^T F)^{-1} F^T x) / s with the
// small core inverse precomputed (rank x rank, fp64).  All arithmetic after the
// fp32 loads is fp64: for x in span(F) the subtraction cancels to a relative size
// s / (s + lambda), so fp32 arithmetic would leave errors of u * lambda / s there
// (0.06 at lambda / s = 1e6), enough to stall PCG.  Both passes stage a tile of F
// (128 rows at rank 64) in LDS with coalesced loads and serve up to 8 rhs from it:
// F^T x gives one fp64 partial per tile (added in index order), the core product
// u = C^-1 t runs once per rhs, and the update re-reads F once more.
//
// RBF: S_ij = exp(-|p_i - p_j|^2 / (2 l^2)) is never stored; each block stages
// 256 points and their x values in LDS and every lane accumulates one row
// (compute bound: one exp per pair).
#include "common.hpp"
#include "ops_internal.hpp"

using namespace ck;

namespace {

constexpr int kMaxRank = 128;
constexpr int kWRhs = 8;        // rhs per pass of the tile kernels
constexpr int kWFloats = 8192;  // LDS floats for one F tile: rows * (rank + 1)

// rows of F per tile: as many as fit 8192 floats of LDS at pitch rank + 1 (128 for rank 63)
__host__ __device__ inline int wb_tile_rows(int k) {
  int r = kWFloats / (k + 1);
  r = r >= 128 ? 128 : (r >= 64 ? 64 : 32);
  return r;
}

// stage rows [i0, i0 + tr) of F (row-major, lda) into Fs[row * (k+1) + c].  Rows past n are
// loaded from row n - 1 (a valid address) and never count: the F^T x pass meets x = 0 there
// and the update pass skips them.  Eight loads per thread go out before the first LDS store.
__device__ __forceinline__ void wb_stage(const float* __restrict__ F, int64_t lda, int64_t n, int k, int64_t i0,
                                         int tr, float* Fs) {
  const int total = tr * k;
  for (int e0 = threadIdx.x; e0 < total; e0 += 8 * 256) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int e = e0 + 256 * j < total ? e0 + 256 * j : total - 1;
      const int64_t i = i0 + e / k;
      v[j] = F[(i < n ? i : n - 1) * lda + e % k];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int e = e0 + 256 * j;
      if (e < total) Fs[(e / k) * (k + 1) + e % k] = v[j];
    }
  }
}

// part[(b*k + c)*P + tile] = sum over the tile's rows of F[i][c] x_b[i] (fp64), for the
// kWRhs rhs starting at b0; F is read once per tile for all of them
__global__ __launch_bounds__(256) void wb_ftx_tiles(int64_t n, int k, const float* __restrict__ F, int64_t lda,
                                                    const float* __restrict__ X, int64_t ldx, int nrhs, int b0,
                                                    double* __restrict__ part, int P, const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ float Fs[kWFloats];
  __shared__ float xs[kWRhs][128];
  const int tr = wb_tile_rows(k);
  const int tile = blockIdx.x;
  const int64_t i0 = (int64_t)tile * tr;
  const int nb = nrhs - b0 < kWRhs ? nrhs - b0 : kWRhs;
  wb_stage(F, lda, n, k, i0, tr, Fs);
  for (int e = threadIdx.x; e < kWRhs * tr; e += 256) {
    const int b = e / tr, r = e % tr;
    const int64_t i = i0 + r;
    xs[b][r] = (b < nb && i < n) ? X[(int64_t)(b0 + b) * ldx + i] : 0.f;
  }
  __syncthreads();
  for (int o = threadIdx.x; o < k * nb; o += 256) {
    const int c = o % k, b = o / k;
    double acc = 0.0;
    for (int r = 0; r < tr; ++r) acc = fma((double)Fs[r * (k + 1) + c], (double)xs[b][r], acc);
    part[((int64_t)(b0 + b) * k + c) * P + tile] = acc;
  }
}

// per rhs b: t = sum of the tile partials (index order), u = Cinv t (both fp64)
// u_b = C^-1 (F^T x_b): the P tile partials of each of the k entries are summed in
// kWbSlices fixed slices of consecutive tiles (each slice in tile order, loads eight at a
// time), the slices then in slice order; then the k x k product, one entry per thread
constexpr int kWbSlices = 1024 / kMaxRank;
__global__ __launch_bounds__(1024) void wb_core(int k, const double* __restrict__ part, int P,
                                                const double* __restrict__ Cinv, double* __restrict__ u,
                                                const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ double red[kWbSlices][kMaxRank];
  __shared__ double ts[kMaxRank];
  const int b = blockIdx.x;
  const int c = threadIdx.x % kMaxRank, sl = threadIdx.x / kMaxRank;
  if (c < k) {
    const int per = (P + kWbSlices - 1) / kWbSlices;
    const int p0 = sl * per, p1 = p0 + per < P ? p0 + per : P;
    const double* pc = part + ((int64_t)b * k + c) * P;
    double acc = 0.0;
    int p = p0;
    for (; p + 8 <= p1; p += 8) {
      double v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = pc[p + q];
#pragma unroll
      for (int q = 0; q < 8; ++q) acc += v[q];
    }
    for (; p < p1; ++p) acc += pc[p];
    red[sl][c] = acc;
  }
  __syncthreads();
  if (threadIdx.x < k) {
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < kWbSlices; ++q) acc += red[q][threadIdx.x];
    ts[threadIdx.x] = acc;
  }
  __syncthreads();
  if (threadIdx.x < k) {
    const int cc = threadIdx.x;
    double acc = 0.0;
    for (int j = 0; j < k; ++j) acc = fma(Cinv[cc * k + j], ts[j], acc);
    u[(int64_t)b * k + cc] = acc;
  }
}

// Y_b = scale (x_b - F u_b) + shift x_b for the tile's rows and kWRhs rhs from b0
__global__ __launch_bounds__(256) void wb_apply_tiles(int64_t n, int k, const float* __restrict__ F, int64_t lda,
                                                      const double* __restrict__ u, const float* __restrict__ X,
                                                      int64_t ldx, float* __restrict__ Y, int64_t ldy, int nrhs,
                                                      int b0, double scale, double shift,
                                                      const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ float Fs[kWFloats];
  __shared__ double us[kWRhs][kMaxRank];
  const int tr = wb_tile_rows(k);
  const int64_t i0 = (int64_t)blockIdx.x * tr;
  const int nb = nrhs - b0 < kWRhs ? nrhs - b0 : kWRhs;
  wb_stage(F, lda, n, k, i0, tr, Fs);
  for (int e = threadIdx.x; e < kWRhs * k; e += 256) {
    const int b = e / k, c = e % k;
    us[b][c] = b < nb ? u[(int64_t)(b0 + b) * k + c] : 0.0;
  }
  __syncthreads();
  for (int o = threadIdx.x; o < tr * nb; o += 256) {
    const int r = o % tr, b = o / tr;
    const int64_t i = i0 + r;
    if (i >= n) continue;
    double acc = 0.0;
    const float* fr = Fs + r * (k + 1);
    for (int c = 0; c < k; ++c) acc = fma((double)fr[c], us[b][c], acc);
    const double x = (double)X[(int64_t)(b0 + b) * ldx + i];
    Y[(int64_t)(b0 + b) * ldy + i] = (float)(scale * (x - acc) + shift * x);
  }
}

constexpr int kRbfTile = 256;
constexpr int kRbfMaxDim = 8;
constexpr int kRbfNB = 8;

// the stationary kernel k(d2) for d2 = |p_i - p_j|^2: FAMILY 0 exp(-d2 / (2 l^2)) (RBF),
// 1 exp(-r / l) (Matern 1/2), 2 (1 + a) exp(-a) with a = sqrt(3) r / l (Matern 3/2),
// 3 (1 + a + a^2 / 3) exp(-a) with a = sqrt(5) r / l (Matern 5/2)
template <int FAMILY>
__device__ __forceinline__ float stationary(float d2, float inv2l2, float inv_l) {
  if (FAMILY == 0) return expf(-d2 * inv2l2);
  const float r = sqrtf(d2);
  if (FAMILY == 1) return expf(-r * inv_l);
  if (FAMILY == 2) {
    const float a = 1.7320508075688772f * r * inv_l;
    return (1.f + a) * expf(-a);
  }
  const float a = 2.23606797749979f * r * inv_l;
  return (1.f + a + a * a * (1.f / 3.f)) * expf(-a);
}

template <int FAMILY>
__global__ __launch_bounds__(256) void rbf_kernel(int n, int dim, const float* __restrict__ P, int64_t ldp,
                                                  float inv2l2, float inv_l, const float* __restrict__ X,
                                                  int64_t ldx, float* __restrict__ Y, int64_t ldy, int nrhs,
                                                  float scale, float shift, const int* __restrict__ done) {
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
      const float kv = stationary<FAMILY>(d2, inv2l2, inv_l);
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

size_t wb_fused_bytes(int64_t n, int nrhs, int k) {
  auto al = [](size_t v) { return (v + 255) / 256 * 256; };
  const size_t T = cdiv(n, kWbFusedRows), np = dense_xy_nparts(n);
  return al(4 * (size_t)nrhs * np) + 2 * al(4 * (size_t)nrhs * T) + al(8 * (size_t)nrhs * k * T) +
         al(8 * (size_t)nrhs * k);
}

size_t woodbury_workspace_bytes(const ck_op* op, int nrhs) {
  const int k = op->rank < 1 ? 1 : op->rank;
  const size_t P = cdiv(op->n, wb_tile_rows(k));
  size_t b = sizeof(double) * ((size_t)nrhs * k * P + (size_t)nrhs * k);
  // room for the PCG step's fused Woodbury passes (cg.hip) as well
  if (nrhs >= 3 && nrhs <= 8 && k <= 64) {
    const size_t f = wb_fused_bytes(op->n, nrhs, k);
    if (f > b) b = f;
  }
  return b;
}

int woodbury_apply(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, void* work,
                   size_t work_bytes, const int* done, hipStream_t s) {
  const int k = op->rank;
  if (k < 1 || k > kMaxRank || !op->a || !op->aux || op->lda < k) return CK_ERR_BAD_ARG;
  if (!work || work_bytes < woodbury_workspace_bytes(op, nrhs)) return CK_ERR_WORKSPACE;
  const int64_t n = op->n;
  const int tr = wb_tile_rows(k);
  const int P = (int)cdiv(n, tr);
  double* part = static_cast<double*>(work);
  double* u = part + (size_t)nrhs * k * P;
  const float* F = static_cast<const float*>(op->a);
  for (int b0 = 0; b0 < nrhs; b0 += kWRhs)
    hipLaunchKernelGGL(wb_ftx_tiles, dim3(P), dim3(256), 0, s, n, k, F, op->lda, X, ldx, nrhs, b0, part, P, done);
  hipLaunchKernelGGL(wb_core, dim3(nrhs), dim3(1024), 0, s, k, (const double*)part, P,
                     reinterpret_cast<const double*>(op->aux), u, done);
  for (int b0 = 0; b0 < nrhs; b0 += kWRhs)
    hipLaunchKernelGGL(wb_apply_tiles, dim3(P), dim3(256), 0, s, n, k, F, op->lda, (const double*)u, X, ldx, Y, ldy,
                       nrhs, b0, (double)op->scale, (double)op->shift, done);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int rbf_apply(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, const int* done,
              hipStream_t s) {
  if (!op->a || op->dim < 1 || op->dim > kRbfMaxDim || op->lda < op->dim || !(op->aux_f > 0.f) || op->rank < 0 ||
      op->rank > 3)
    return CK_ERR_BAD_ARG;
  const float inv2l2 = 1.f / (2.f * op->aux_f * op->aux_f), inv_l = 1.f / op->aux_f;
  const dim3 grid((unsigned)cdiv(op->n, 256), (unsigned)cdiv(nrhs, kRbfNB));
  const float* P = static_cast<const float*>(op->a);
#define CK_RBF(F_)                                                                                               \
  hipLaunchKernelGGL(rbf_kernel<F_>, grid, dim3(256), 0, s, op->n, op->dim, P, op->lda, inv2l2, inv_l, X, ldx, Y, \
                     ldy, nrhs, op->scale, op->shift, done)
  switch (op->rank) {
    case 0:
      CK_RBF(0);
      break;
    case 1:
      CK_RBF(1);
      break;
    case 2:
      CK_RBF(2);
      break;
    default:
      CK_RBF(3);
      break;
  }
#undef CK_RBF
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // namespace ck
