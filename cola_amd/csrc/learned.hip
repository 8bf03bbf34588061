// Learned low-rank preconditioner: a per-point MLP (din -> 128 -> 128 -> dout, GELU),
// its training kernels, and the small dense algebra that turns its output into a
// Woodbury factor; plus the rank-r pivoted Cholesky comparator and the dense RBF
// kernel-matrix builder.
//
//   mlp_fwd      one block = 64 points; the three layers run back to back with the
//                hidden activations in LDS (bank-padded pitch 129), optional
//                pre-activation save for the backward pass    [VALU, ~50k flop/point]
//   mlp_bwd      one block = 64 points; per-block partial parameter gradients in the
//                flat parameter layout, summed over blocks in index order (FP64)
//   gram         C = X^T Y for tall-skinny X, Y (FP64 accumulation, fixed order)
//   rows_gemm    Y = alpha X W + beta Y for tall-skinny X and a small W (LDS-staged)
//   spd_inv      batched Cholesky of small SPD matrices (k <= 64) with jitter,
//                triangular inverse, full inverse                          [latency]
//   energy       the captured-energy loss and its two small gradient factors
//   pivchol      greedy (largest residual diagonal) pivoted Cholesky, Harbrecht,
//                Peters & Schneider (2012): one column kernel per rank step with a
//                fused block argmax of the updated diagonal              [HBM]
//   rbf_dense    K_ij = exp(-|p_i - p_j|^2 / (2 l^2)), exactly symmetric,
//                optional bf16 copy                                         [HBM write]
//   dense_f64acc Y = scale A X + shift X with fp32 A and fp64 vectors/accumulation
//                (refinement residuals of dense systems)                   [HBM]
#include "common.hpp"

using namespace ck;

namespace {

constexpr int kH = 128;     // hidden width
constexpr int kMaxIn = 8;   // input dimension limit
constexpr int kMaxOut = 128;
constexpr int kRows = 64;   // points per block
constexpr int kPitch = kH + 1;

struct MlpLayout {
  int64_t w1, b1, w2, b2, w3, b3, total;
};

__host__ __device__ inline MlpLayout mlp_layout(int din, int dout) {
  MlpLayout L;
  L.w1 = 0;
  L.b1 = L.w1 + (int64_t)din * kH;
  L.w2 = L.b1 + kH;
  L.b2 = L.w2 + (int64_t)kH * kH;
  L.w3 = L.b2 + kH;
  L.b3 = L.w3 + (int64_t)kH * dout;
  L.total = L.b3 + dout;
  return L;
}

// exact (erf) GELU and its derivative Phi(z) + z phi(z)
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  const float cdf = 0.5f * (1.f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return fmaf(z, pdf, cdf);
}

// ------------------------------------------------------------------ MLP forward
template <bool SAVE>
__global__ __launch_bounds__(256) void mlp_fwd_kernel(int n, int din, int dout, const float* __restrict__ X,
                                                      int64_t ldx, const float* __restrict__ prm,
                                                      float* __restrict__ G, int64_t gsr, int64_t gsc,
                                                      float* __restrict__ Z1, float* __restrict__ Z2, int64_t ldz) {
  __shared__ float xs[kRows][kMaxIn];
  __shared__ float h1[kRows][kPitch];
  __shared__ float h2[kRows][kPitch];
  const MlpLayout L = mlp_layout(din, dout);
  // g is the wave index: made wave-uniform (an SGPR), so the weight loads below, whose
  // addresses depend only on g and the loop counters, become scalar loads
  const int r = threadIdx.x & 63, g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t row = (int64_t)blockIdx.x * kRows + r;
  const bool valid = row < n;
  for (int e = threadIdx.x; e < kRows * kMaxIn; e += 256) {
    const int rr = e / kMaxIn, d = e % kMaxIn;
    const int64_t i = (int64_t)blockIdx.x * kRows + rr;
    xs[rr][d] = (i < n && d < din) ? X[i * ldx + d] : 0.f;
  }
  __syncthreads();
  const int c0 = g * 32;  // each wave owns 32 hidden units
  {
    float acc[32];
#pragma unroll
    for (int cc = 0; cc < 32; ++cc) acc[cc] = prm[L.b1 + c0 + cc];
    for (int d = 0; d < din; ++d) {
      const float xv = xs[r][d];
      const float* w = prm + L.w1 + (int64_t)d * kH + c0;
#pragma unroll
      for (int cc = 0; cc < 32; ++cc) acc[cc] = fmaf(xv, w[cc], acc[cc]);
    }
#pragma unroll
    for (int cc = 0; cc < 32; ++cc) {
      if (SAVE && valid) Z1[(int64_t)(c0 + cc) * ldz + row] = acc[cc];
      h1[r][c0 + cc] = gelu(acc[cc]);
    }
  }
  __syncthreads();
  {
    float acc[32];
#pragma unroll
    for (int cc = 0; cc < 32; ++cc) acc[cc] = prm[L.b2 + c0 + cc];
    for (int j = 0; j < kH; ++j) {
      const float hv = h1[r][j];
      const float* w = prm + L.w2 + (int64_t)j * kH + c0;
#pragma unroll
      for (int cc = 0; cc < 32; ++cc) acc[cc] = fmaf(hv, w[cc], acc[cc]);
    }
#pragma unroll
    for (int cc = 0; cc < 32; ++cc) {
      if (SAVE && valid) Z2[(int64_t)(c0 + cc) * ldz + row] = acc[cc];
      h2[r][c0 + cc] = gelu(acc[cc]);
    }
  }
  __syncthreads();
  {
    const int per = (dout + 3) / 4;  // <= 32 outputs per wave
    const int o0 = g * per;
    float acc[32];
#pragma unroll
    for (int cc = 0; cc < 32; ++cc) acc[cc] = (cc < per && o0 + cc < dout) ? prm[L.b3 + o0 + cc] : 0.f;
    for (int j = 0; j < kH; ++j) {
      const float hv = h2[r][j];
      const float* w = prm + L.w3 + (int64_t)j * dout + o0;
#pragma unroll
      for (int cc = 0; cc < 32; ++cc)
        if (cc < per && o0 + cc < dout) acc[cc] = fmaf(hv, w[cc], acc[cc]);
    }
    if (valid) {
#pragma unroll
      for (int cc = 0; cc < 32; ++cc)
        if (cc < per && o0 + cc < dout) G[row * gsr + (int64_t)(o0 + cc) * gsc] = acc[cc];
    }
  }
}

// ----------------------------------------------------------------- MLP backward
// Given dL/dG for a block of 64 points, writes that block's contribution to every
// parameter gradient (flat layout) into part[blockIdx.x * total + ...].
__global__ __launch_bounds__(256) void mlp_bwd_kernel(int n, int din, int dout, const float* __restrict__ X,
                                                      int64_t ldx, const float* __restrict__ prm,
                                                      const float* __restrict__ Z1, const float* __restrict__ Z2,
                                                      int64_t ldz, const float* __restrict__ dG, int64_t dsr,
                                                      int64_t dsc, float* __restrict__ part) {
  __shared__ float xs[kRows][kMaxIn];
  __shared__ float a[kRows][kPitch];   // h2, then h1, then dZ1
  __shared__ float dz[kRows][kPitch];  // dZ2
  __shared__ float dg[kRows][kPitch];  // dG (dout <= 128)
  const MlpLayout L = mlp_layout(din, dout);
  float* out = part + (int64_t)blockIdx.x * L.total;
  const int t = threadIdx.x;
  const int r = t & 63, g = __builtin_amdgcn_readfirstlane(t >> 6);  // wave-uniform: scalar weight loads
  const int64_t row = (int64_t)blockIdx.x * kRows + r;
  const bool valid = row < n;
  for (int e = t; e < kRows * kMaxIn; e += 256) {
    const int rr = e / kMaxIn, d = e % kMaxIn;
    const int64_t i = (int64_t)blockIdx.x * kRows + rr;
    xs[rr][d] = (i < n && d < din) ? X[i * ldx + d] : 0.f;
  }
  for (int c = g; c < dout; c += 4) dg[r][c] = valid ? dG[row * dsr + (int64_t)c * dsc] : 0.f;
  for (int j = g; j < kH; j += 4) a[r][j] = valid ? gelu(Z2[(int64_t)j * ldz + row]) : 0.f;
  __syncthreads();

  // dW3[j][c] = sum_r h2[r][j] dG[r][c]; db3[c] = sum_r dG[r][c]
  {
    const int j = t & 127, par = t >> 7;
    const int nq = (dout - par + 1) / 2;  // columns c = par + 2 q
    for (int q0 = 0; q0 < nq; q0 += 16) {
      float acc[16];
#pragma unroll
      for (int qq = 0; qq < 16; ++qq) acc[qq] = 0.f;
      for (int rr = 0; rr < kRows; ++rr) {
        const float av = a[rr][j];
#pragma unroll
        for (int qq = 0; qq < 16; ++qq) {
          const int c = par + 2 * (q0 + qq);
          if (q0 + qq < nq) acc[qq] = fmaf(av, dg[rr][c], acc[qq]);
        }
      }
#pragma unroll
      for (int qq = 0; qq < 16; ++qq)
        if (q0 + qq < nq) out[L.w3 + (int64_t)j * dout + par + 2 * (q0 + qq)] = acc[qq];
    }
    for (int c = t; c < dout; c += 256) {
      float s = 0.f;
      for (int rr = 0; rr < kRows; ++rr) s += dg[rr][c];
      out[L.b3 + c] = s;
    }
  }
  // dZ2[r][j] = (sum_c dG[r][c] W3[j][c]) * gelu'(Z2[r][j])
  for (int jj = 0; jj < 32; ++jj) {
    const int j = g * 32 + jj;
    const float* w = prm + L.w3 + (int64_t)j * dout;
    float acc = 0.f;
    for (int c = 0; c < dout; ++c) acc = fmaf(dg[r][c], w[c], acc);
    dz[r][j] = valid ? acc * gelu_grad(Z2[(int64_t)j * ldz + row]) : 0.f;
  }
  __syncthreads();
  for (int j = g; j < kH; j += 4) a[r][j] = valid ? gelu(Z1[(int64_t)j * ldz + row]) : 0.f;
  __syncthreads();
  // dW2[j][k] = sum_r h1[r][j] dZ2[r][k]; db2[k] = sum_r dZ2[r][k]
  {
    const int j = t & 127, par = t >> 7;
    for (int q0 = 0; q0 < kH / 2; q0 += 16) {
      float acc[16];
#pragma unroll
      for (int qq = 0; qq < 16; ++qq) acc[qq] = 0.f;
      for (int rr = 0; rr < kRows; ++rr) {
        const float av = a[rr][j];
#pragma unroll
        for (int qq = 0; qq < 16; ++qq) acc[qq] = fmaf(av, dz[rr][par + 2 * (q0 + qq)], acc[qq]);
      }
#pragma unroll
      for (int qq = 0; qq < 16; ++qq) out[L.w2 + (int64_t)j * kH + par + 2 * (q0 + qq)] = acc[qq];
    }
    for (int k = t; k < kH; k += 256) {
      float s = 0.f;
      for (int rr = 0; rr < kRows; ++rr) s += dz[rr][k];
      out[L.b2 + k] = s;
    }
  }
  // dZ1[r][j] = (sum_k dZ2[r][k] W2[j][k]) * gelu'(Z1[r][j]), kept in registers until a is free
  float d1[32];
  for (int jj = 0; jj < 32; ++jj) {
    const int j = g * 32 + jj;
    const float* w = prm + L.w2 + (int64_t)j * kH;
    float acc = 0.f;
    for (int k = 0; k < kH; ++k) acc = fmaf(dz[r][k], w[k], acc);
    d1[jj] = valid ? acc * gelu_grad(Z1[(int64_t)j * ldz + row]) : 0.f;
  }
  __syncthreads();
#pragma unroll
  for (int jj = 0; jj < 32; ++jj) a[r][g * 32 + jj] = d1[jj];
  __syncthreads();
  // dW1[d][j] = sum_r x[r][d] dZ1[r][j]; db1[j] = sum_r dZ1[r][j]
  {
    const int j = t & 127, par = t >> 7;
    for (int d = par; d < din; d += 2) {
      float acc = 0.f;
      for (int rr = 0; rr < kRows; ++rr) acc = fmaf(xs[rr][d], a[rr][j], acc);
      out[L.w1 + (int64_t)d * kH + j] = acc;
    }
    if (t < kH) {
      float s = 0.f;
      for (int rr = 0; rr < kRows; ++rr) s += a[rr][t];
      out[L.b1 + t] = s;
    }
  }
}

// grad[p] = scale * sum_b part[b][p] (blocks in index order, FP64)
__global__ __launch_bounds__(256) void grad_reduce_kernel(int nblk, int64_t total, const float* __restrict__ part,
                                                          float scale, float* __restrict__ grad) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= total) return;
  double acc = 0.0;
  for (int b = 0; b < nblk; ++b) acc += (double)part[(int64_t)b * total + p];
  grad[p] = (float)(scale * acc);
}

// torch.optim.Adam update (no weight decay): m, v moments; bc1 = 1 - b1^t, bc2 = 1 - b2^t
__global__ __launch_bounds__(256) void adam_kernel(int64_t count, float* __restrict__ p, const float* __restrict__ gr,
                                                   float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                                   float b2, float eps, float bc1, float bc2) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  const float gi = gr[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float denom = sqrtf(vi) / sqrtf(bc2) + eps;
  p[i] -= (lr / bc1) * (mi / denom);
}

// ------------------------------------------------------------------ gram C = X^T Y
constexpr int kGramRows = 512;

__global__ __launch_bounds__(256) void gram_partials(int64_t n, int ka, int kb, const float* __restrict__ X,
                                                     int64_t xsr, int64_t xsc, const float* __restrict__ Y,
                                                     int64_t ysr, int64_t ysc, double* __restrict__ part) {
  __shared__ float xs[64][33];
  __shared__ float ys[64][33];
  const int ta = blockIdx.y, tb = blockIdx.z;
  const int al = threadIdx.x >> 3, bq = (threadIdx.x & 7) * 4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int64_t r0 = (int64_t)blockIdx.x * kGramRows;
  const int64_t r1 = r0 + kGramRows < n ? r0 + kGramRows : n;
  for (int64_t rb = r0; rb < r1; rb += 64) {
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 32; e += 256) {
      const int rr = e & 63, cc = e >> 6;
      const int64_t i = rb + rr;
      const int ai = ta * 32 + cc, bi = tb * 32 + cc;
      xs[rr][cc] = (i < r1 && ai < ka) ? X[i * xsr + (int64_t)ai * xsc] : 0.f;
      ys[rr][cc] = (i < r1 && bi < kb) ? Y[i * ysr + (int64_t)bi * ysc] : 0.f;
    }
    __syncthreads();
    for (int rr = 0; rr < 64; ++rr) {
      const double xv = (double)xs[rr][al];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = fma(xv, (double)ys[rr][bq + q], acc[q]);
    }
  }
  const int ai = ta * 32 + al;
  if (ai >= ka) return;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int bi = tb * 32 + bq + q;
    if (bi < kb) part[(int64_t)blockIdx.x * ka * kb + (int64_t)ai * kb + bi] = acc[q];
  }
}

__global__ __launch_bounds__(256) void gram_reduce(int nchunks, int64_t count, const double* __restrict__ part,
                                                   double* __restrict__ C) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= count) return;
  double acc = 0.0;
  for (int c = 0; c < nchunks; ++c) acc += part[(int64_t)c * count + idx];
  C[idx] = acc;
}

// --------------------------------------------------------- rows GEMM Y = a X W + b Y
constexpr int kGemmMax = 128;

__global__ __launch_bounds__(256) void rows_gemm_kernel(int64_t n, int kin, int kout, const float* __restrict__ X,
                                                        int64_t xsr, int64_t xsc, const float* __restrict__ W,
                                                        float alpha, float beta, float* __restrict__ Y, int64_t ysr,
                                                        int64_t ysc) {
  __shared__ float ws[kGemmMax * 32];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int c0 = blockIdx.y * 32;
  const int nc = kout - c0 < 32 ? kout - c0 : 32;
  for (int e = threadIdx.x; e < kin * 32; e += 256) {
    const int j = e / 32, cc = e % 32;
    ws[e] = cc < nc ? W[(int64_t)j * kout + c0 + cc] : 0.f;
  }
  __syncthreads();
  if (i >= n) return;
  float acc[32];
#pragma unroll
  for (int cc = 0; cc < 32; ++cc) acc[cc] = 0.f;
  for (int j = 0; j < kin; ++j) {
    const float xv = X[i * xsr + (int64_t)j * xsc];
#pragma unroll
    for (int cc = 0; cc < 32; ++cc) acc[cc] = fmaf(xv, ws[j * 32 + cc], acc[cc]);
  }
#pragma unroll
  for (int cc = 0; cc < 32; ++cc) {
    if (cc < nc) {
      float* y = Y + i * ysr + (int64_t)(c0 + cc) * ysc;
      *y = beta == 0.f ? alpha * acc[cc] : fmaf(beta, *y, alpha * acc[cc]);
    }
  }
}

// ------------------------------------------------------ small SPD: chol, L^-1, A^-1
constexpr int kSpdMax = 64;
constexpr int kPack = kSpdMax * (kSpdMax + 1) / 2;  // packed lower triangle

__device__ __forceinline__ int pk(int i, int j) { return i * (i + 1) / 2 + j; }  // i >= j

// One block per matrix, everything in LDS (packed lower triangles of L and X = L^-1):
// A + (jitter_abs + jitter_rel * trace(A) / k) I = L L^T.
// linv (required, k x k row-major lower) = L^-1; inv64 / inv32 = A^-1; linvT32 = L^-T.
// status[b] = 0, or j + 1 when the j-th pivot is not positive.
__global__ __launch_bounds__(256) void spd_inv_kernel(int k, const double* __restrict__ A, int64_t lda, int64_t sa,
                                                      double jitter_rel, double jitter_abs, double* __restrict__ linv,
                                                      double* __restrict__ inv64, float* __restrict__ inv32,
                                                      float* __restrict__ linvT32, int* __restrict__ status) {
  __shared__ double Ls[kPack];
  __shared__ double Xs[kPack];
  __shared__ double shift;
  __shared__ int bad;
  const int b = blockIdx.x, t = threadIdx.x;
  const double* Ab = A + (int64_t)b * sa;
  for (int e = t; e < k * k; e += 256) {
    const int i = e / k, j = e % k;
    if (j <= i) Ls[pk(i, j)] = Ab[(int64_t)i * lda + j];
  }
  __syncthreads();
  if (t == 0) {
    double tr = 0.0;
    for (int i = 0; i < k; ++i) tr += Ls[pk(i, i)];
    shift = jitter_abs + jitter_rel * tr / k;
    bad = 0;
  }
  __syncthreads();
  if (t < k) Ls[pk(t, t)] += shift;
  for (int j = 0; j < k; ++j) {
    __syncthreads();
    if (t == 0) {
      const double v = Ls[pk(j, j)];
      if (!(v > 0.0)) {
        bad = j + 1;
      } else {
        Ls[pk(j, j)] = sqrt(v);
      }
    }
    __syncthreads();
    if (bad) break;
    const double djj = Ls[pk(j, j)];
    for (int i = j + 1 + t; i < k; i += 256) Ls[pk(i, j)] /= djj;
    __syncthreads();
    // trailing update of the lower triangle: row i = j+1 + t/4, every 4th column from j+1+t%4
    {
      const int i = j + 1 + (t >> 2);
      if (i < k) {
        const double lij = Ls[pk(i, j)];
        const int base = i * (i + 1) / 2;
        for (int l = j + 1 + (t & 3); l <= i; l += 4) Ls[base + l] -= lij * Ls[pk(l, j)];
      }
    }
  }
  __syncthreads();
  if (t == 0) status[b] = bad;
  if (bad) return;
  // column c of X = L^-1 by forward substitution (one thread per column)
  if (t < k) {
    const int c = t;
    Xs[pk(c, c)] = 1.0 / Ls[pk(c, c)];
    for (int i = c + 1; i < k; ++i) {
      double acc = 0.0;
      for (int q = c; q < i; ++q) acc = fma(Ls[pk(i, q)], Xs[pk(q, c)], acc);
      Xs[pk(i, c)] = -acc / Ls[pk(i, i)];
    }
  }
  __syncthreads();
  double* Li = linv + (int64_t)b * k * k;
  for (int e = t; e < k * k; e += 256) {
    const int r = e / k, c = e % k;
    Li[e] = c <= r ? Xs[pk(r, c)] : 0.0;
    if (linvT32) linvT32[(int64_t)b * k * k + e] = r <= c ? (float)Xs[pk(c, r)] : 0.f;
    if (inv64 || inv32) {
      const int c0 = r > c ? r : c;
      double acc = 0.0;
      for (int q = c0; q < k; ++q) acc = fma(Xs[pk(q, r)], Xs[pk(q, c)], acc);
      if (inv64) inv64[(int64_t)b * k * k + e] = acc;
      if (inv32) inv32[(int64_t)b * k * k + e] = (float)acc;
    }
  }
}

// ------------------------------------------------------------- captured-energy loss
// With S^-1 (k x k), T = U^T G (m x k) and weights w (m): P = T S^-1,
//   f  = sum_i w_i sum_a P[i][a] T[i][a] / sum w      (= tr(S^-1 G^T U W U^T G) / sum w)
//   C1 = -(2 / sum w) diag(w) P          (m x k)  so that  dL/dG = U C1 + G C2
//   C2 =  (2 / sum w) P^T diag(w) P      (k x k)  for the loss L = 1 - f
// T and the lower triangle of S^-1 live in LDS; P overwrites T once every thread has
// read its row.
__global__ __launch_bounds__(256) void energy_kernel(int k, int m, const double* __restrict__ Sinv,
                                                     const double* __restrict__ T, const double* __restrict__ w,
                                                     float* __restrict__ C1, float* __restrict__ C2,
                                                     double* __restrict__ loss) {
  __shared__ double Ts[kSpdMax * kSpdMax];
  __shared__ double Sp[kPack];
  __shared__ double ws[kSpdMax];
  __shared__ double red[4];
  const int t = threadIdx.x;
  for (int e = t; e < m * k; e += 256) Ts[e] = T[e];
  for (int e = t; e < k * k; e += 256) {
    const int i = e / k, j = e % k;
    if (j <= i) Sp[pk(i, j)] = Sinv[e];
  }
  for (int i = t; i < m; i += 256) ws[i] = w[i];
  __syncthreads();
  double wsum = 0.0;
  for (int i = 0; i < m; ++i) wsum += ws[i];
  const double c = 2.0 / wsum;
  const int i = t >> 2, g = t & 3;
  double pv[kSpdMax / 4];
  double f = 0.0;
  if (i < m) {
#pragma unroll
    for (int u = 0; u < kSpdMax / 4; ++u) {
      const int a = g + 4 * u;
      pv[u] = 0.0;
      if (a < k) {
        double acc = 0.0;
        for (int q = 0; q < k; ++q) acc = fma(Ts[i * k + q], Sp[q >= a ? pk(q, a) : pk(a, q)], acc);
        pv[u] = acc;
        C1[i * k + a] = (float)(-c * ws[i] * acc);
        f = fma(ws[i] * acc, Ts[i * k + a], f);
      }
    }
  }
  f = block_sum<256>(f, red);  // its barriers also order every read of Ts before the writes below
  if (i < m) {
#pragma unroll
    for (int u = 0; u < kSpdMax / 4; ++u) {
      const int a = g + 4 * u;
      if (a < k) Ts[i * k + a] = pv[u];
    }
  }
  __syncthreads();
  for (int e = t; e < k * k; e += 256) {
    const int a = e / k, bb = e % k;
    double acc = 0.0;
    for (int r = 0; r < m; ++r) acc = fma(ws[r] * Ts[r * k + a], Ts[r * k + bb], acc);
    C2[e] = (float)(c * acc);
  }
  if (t == 0) loss[0] = 1.0 - f / wsum;
}

// ------------------------------------------------------------ pivoted Cholesky
constexpr int kPcRows = 1024;  // rows per block (4 per thread)

__device__ __forceinline__ bool pc_better(float v, int i, float bv, int bi) {
  return v > bv || (v == bv && i < bi);
}

// block argmax of (value, index) with ties to the smaller index; order independent
__device__ __forceinline__ void pc_block_argmax(float v, int i, float* sv, int* si, float* outv, int* outi) {
  for (int off = 32; off >= 1; off >>= 1) {
    const float ov = __shfl_xor(v, off);
    const int oi = __shfl_xor(i, off);
    if (pc_better(ov, oi, v, i)) {
      v = ov;
      i = oi;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sv[wave] = v;
    si[wave] = i;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float bv = sv[0];
    int bi = si[0];
    for (int q = 1; q < 4; ++q)
      if (pc_better(sv[q], si[q], bv, bi)) {
        bv = sv[q];
        bi = si[q];
      }
    *outv = bv;
    *outi = bi;
  }
}

__global__ __launch_bounds__(256) void pc_init_kernel(int n, const float* __restrict__ A, int64_t lda,
                                                      float* __restrict__ d, float* __restrict__ pval,
                                                      int* __restrict__ pidx) {
  __shared__ float sv[4];
  __shared__ int si[4];
  float bv = -INFINITY;
  int bi = 0x7FFFFFFF;
  for (int q = 0; q < kPcRows / 256; ++q) {
    const int i = blockIdx.x * kPcRows + q * 256 + threadIdx.x;
    if (i < n) {
      const float v = A[(int64_t)i * lda + i];
      d[i] = v;
      if (pc_better(v, i, bv, bi)) {
        bv = v;
        bi = i;
      }
    }
  }
  pc_block_argmax(bv, bi, sv, si, pval + blockIdx.x, pidx + blockIdx.x);
}

// pivot j = argmax over the block partials (one wave, fixed order)
__global__ __launch_bounds__(64) void pc_select_kernel(int nparts, const float* __restrict__ pval,
                                                       const int* __restrict__ pidx, int j, int* __restrict__ piv,
                                                       float* __restrict__ dpiv) {
  float bv = -INFINITY;
  int bi = 0x7FFFFFFF;
  for (int p = threadIdx.x; p < nparts; p += 64)
    if (pc_better(pval[p], pidx[p], bv, bi)) {
      bv = pval[p];
      bi = pidx[p];
    }
  for (int off = 32; off >= 1; off >>= 1) {
    const float ov = __shfl_xor(bv, off);
    const int oi = __shfl_xor(bi, off);
    if (pc_better(ov, oi, bv, bi)) {
      bv = ov;
      bi = oi;
    }
  }
  if (threadIdx.x == 0) {
    piv[j] = bi;
    dpiv[j] = bv;
  }
}

// column j: l_i = (A[p][i] - sum_{q<j} L_q[i] L_q[p]) / sqrt(d_p); d_i -= l_i^2; d_p = 0.
// A pivot with d_p <= 0 (exhausted rank) gives a zero column.
__global__ __launch_bounds__(256) void pc_column_kernel(int n, const float* __restrict__ A, int64_t lda, int j,
                                                        const int* __restrict__ piv,
                                                        const float* __restrict__ dpiv, float* __restrict__ L,
                                                        int64_t ldl, float* __restrict__ d, float* __restrict__ pval,
                                                        int* __restrict__ pidx) {
  __shared__ float sv[4];
  __shared__ int si[4];
  const int p = piv[j];
  const float dp = dpiv[j];
  const bool ok = dp > 0.f && p >= 0 && p < n;
  const float sq = ok ? sqrtf(dp) : 1.f;
  float bv = -INFINITY;
  int bi = 0x7FFFFFFF;
  for (int q = 0; q < kPcRows / 256; ++q) {
    const int i = blockIdx.x * kPcRows + q * 256 + threadIdx.x;
    if (i < n) {
      float l = 0.f;
      if (ok) {
        float acc = A[(int64_t)p * lda + i];
        for (int c = 0; c < j; ++c) acc = fmaf(-L[(int64_t)c * ldl + i], L[(int64_t)c * ldl + p], acc);
        l = acc / sq;
      }
      L[(int64_t)j * ldl + i] = l;
      const float nd = (ok && i == p) ? 0.f : fmaf(-l, l, d[i]);
      d[i] = nd;
      if (pc_better(nd, i, bv, bi)) {
        bv = nd;
        bi = i;
      }
    }
  }
  pc_block_argmax(bv, bi, sv, si, pval + blockIdx.x, pidx + blockIdx.x);
}

// ------------------------------------------------------------------ RBF dense build
__global__ __launch_bounds__(256) void rbf_dense_kernel(int n, int dim, const float* __restrict__ P, int64_t ldp,
                                                        float inv2l2, float* __restrict__ K, int64_t ldk,
                                                        uint16_t* __restrict__ Kb, int64_t ldkb) {
  const int64_t i = blockIdx.x;
  float pi[kMaxIn];
#pragma unroll
  for (int d = 0; d < kMaxIn; ++d) pi[d] = d < dim ? P[i * ldp + d] : 0.f;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t j = (int64_t)blockIdx.y * 1024 + q * 256 + threadIdx.x;
    if (j < n) {
      float d2 = 0.f;
#pragma unroll
      for (int d = 0; d < kMaxIn; ++d) {
        if (d < dim) {
          const float df = pi[d] - P[j * ldp + d];
          d2 = fmaf(df, df, d2);
        }
      }
      const float v = expf(-d2 * inv2l2);
      if (K) K[i * ldk + j] = v;
      if (Kb) {
        const uint32_t u = __float_as_uint(v);
        Kb[i * ldkb + j] = (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
      }
    }
  }
}

// ------------------------------------------------------- dense apply, fp64 vectors
constexpr int kF64Rows = 8;  // rows per wave
constexpr int kF64NB = 8;    // rhs per pass

__global__ __launch_bounds__(256) void dense_f64acc_kernel(int n, const float* __restrict__ A, int64_t lda,
                                                           const double* __restrict__ X, int64_t ldx,
                                                           double* __restrict__ Y, int64_t ldy, int nb, double scale,
                                                           double shift) {
  __shared__ double xs[kF64NB][256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row0 = blockIdx.x * (4 * kF64Rows) + wave * kF64Rows;
  double acc[kF64Rows][kF64NB];
#pragma unroll
  for (int r = 0; r < kF64Rows; ++r)
#pragma unroll
    for (int b = 0; b < kF64NB; ++b) acc[r][b] = 0.0;
  for (int j0 = 0; j0 < n; j0 += 256) {
    __syncthreads();
    for (int e = threadIdx.x; e < kF64NB * 256; e += 256) {
      const int b = e / 256, jj = e % 256;
      xs[b][jj] = (b < nb && j0 + jj < n) ? X[(int64_t)b * ldx + j0 + jj] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kF64Rows; ++r) {
      const int row = row0 + r;
      if (row >= n) break;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int jj = q * 64 + lane;
        const int j = j0 + jj;
        const double a = j < n ? (double)A[(int64_t)row * lda + j] : 0.0;
#pragma unroll
        for (int b = 0; b < kF64NB; ++b) acc[r][b] = fma(a, xs[b][jj], acc[r][b]);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kF64Rows; ++r) {
#pragma unroll
    for (int b = 0; b < kF64NB; ++b) {
      const double s = wave_sum(acc[r][b]);
      const int row = row0 + r;
      if (lane == 0 && row < n && b < nb) Y[(int64_t)b * ldy + row] = scale * s + shift * X[(int64_t)b * ldx + row];
    }
  }
}

}  // namespace

extern "C" {

int64_t ck_mlp_param_count(int32_t din, int32_t dout) {
  if (din < 1 || din > kMaxIn || dout < 1 || dout > kMaxOut) return -1;
  return mlp_layout(din, dout).total;
}

size_t ck_mlp_grad_workspace_bytes(int64_t n, int32_t din, int32_t dout) {
  if (n < 1 || din < 1 || din > kMaxIn || dout < 1 || dout > kMaxOut) return 0;
  return sizeof(float) * (size_t)cdiv(n, kRows) * (size_t)mlp_layout(din, dout).total;
}

int ck_mlp_forward(int64_t n, int32_t din, int32_t dout, const float* X, int64_t ldx, const float* params, float* G,
                   int64_t gsr, int64_t gsc, float* Z1, float* Z2, int64_t ldz, void* stream) {
  if (n < 1 || n > 0x7FFFFFFF || din < 1 || din > kMaxIn || dout < 1 || dout > kMaxOut || !X || !params || !G ||
      ldx < din)
    return CK_ERR_BAD_ARG;
  if ((Z1 == nullptr) != (Z2 == nullptr) || (Z1 && ldz < n)) return CK_ERR_BAD_ARG;
  const unsigned grid = (unsigned)cdiv(n, kRows);
  hipStream_t s = (hipStream_t)stream;
  if (Z1)
    hipLaunchKernelGGL(mlp_fwd_kernel<true>, dim3(grid), dim3(256), 0, s, (int)n, din, dout, X, ldx, params, G, gsr,
                       gsc, Z1, Z2, ldz);
  else
    hipLaunchKernelGGL(mlp_fwd_kernel<false>, dim3(grid), dim3(256), 0, s, (int)n, din, dout, X, ldx, params, G, gsr,
                       gsc, Z1, Z2, ldz);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_mlp_backward(int64_t n, int32_t din, int32_t dout, const float* X, int64_t ldx, const float* params,
                    const float* Z1, const float* Z2, int64_t ldz, const float* dG, int64_t dsr, int64_t dsc,
                    float grad_scale, float* grad, void* work, size_t work_bytes, void* stream) {
  if (n < 1 || n > 0x7FFFFFFF || din < 1 || din > kMaxIn || dout < 1 || dout > kMaxOut || !X || !params || !Z1 ||
      !Z2 || !dG || !grad || ldx < din || ldz < n)
    return CK_ERR_BAD_ARG;
  if (!work || work_bytes < ck_mlp_grad_workspace_bytes(n, din, dout)) return CK_ERR_WORKSPACE;
  const int nblk = (int)cdiv(n, kRows);
  const int64_t total = mlp_layout(din, dout).total;
  hipStream_t s = (hipStream_t)stream;
  float* part = static_cast<float*>(work);
  hipLaunchKernelGGL(mlp_bwd_kernel, dim3(nblk), dim3(256), 0, s, (int)n, din, dout, X, ldx, params, Z1, Z2, ldz, dG,
                     dsr, dsc, part);
  hipLaunchKernelGGL(grad_reduce_kernel, dim3((unsigned)cdiv(total, 256)), dim3(256), 0, s, nblk, total,
                     (const float*)part, grad_scale, grad);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_adam_step(int64_t count, float* params, const float* grad, float* m, float* v, float lr, float beta1,
                 float beta2, float eps, int32_t step, void* stream) {
  if (count < 1 || !params || !grad || !m || !v || step < 1) return CK_ERR_BAD_ARG;
  const float bc1 = 1.f - powf(beta1, (float)step);
  const float bc2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)cdiv(count, 256)), dim3(256), 0, (hipStream_t)stream, count, params,
                     grad, m, v, lr, beta1, beta2, eps, bc1, bc2);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

size_t ck_gram_workspace_bytes(int64_t n, int32_t ka, int32_t kb) {
  if (n < 1 || ka < 1 || kb < 1) return 0;
  return sizeof(double) * (size_t)cdiv(n, kGramRows) * (size_t)ka * (size_t)kb;
}

int ck_gram_f64(int64_t n, int32_t ka, int32_t kb, const float* X, int64_t xsr, int64_t xsc, const float* Y,
                int64_t ysr, int64_t ysc, double* C, void* work, size_t work_bytes, void* stream) {
  if (n < 1 || ka < 1 || kb < 1 || !X || !Y || !C) return CK_ERR_BAD_ARG;
  if (!work || work_bytes < ck_gram_workspace_bytes(n, ka, kb)) return CK_ERR_WORKSPACE;
  const int nch = (int)cdiv(n, kGramRows);
  hipStream_t s = (hipStream_t)stream;
  double* part = static_cast<double*>(work);
  hipLaunchKernelGGL(gram_partials, dim3(nch, (unsigned)cdiv(ka, 32), (unsigned)cdiv(kb, 32)), dim3(256), 0, s, n, ka,
                     kb, X, xsr, xsc, Y, ysr, ysc, part);
  const int64_t count = (int64_t)ka * kb;
  hipLaunchKernelGGL(gram_reduce, dim3((unsigned)cdiv(count, 256)), dim3(256), 0, s, nch, count, (const double*)part,
                     C);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_rows_gemm_f32(int64_t n, int32_t kin, int32_t kout, const float* X, int64_t xsr, int64_t xsc, const float* W,
                     float alpha, float beta, float* Y, int64_t ysr, int64_t ysc, void* stream) {
  if (n < 1 || kin < 1 || kin > kGemmMax || kout < 1 || !X || !W || !Y) return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(rows_gemm_kernel, dim3((unsigned)cdiv(n, 256), (unsigned)cdiv(kout, 32)), dim3(256), 0,
                     (hipStream_t)stream, n, kin, kout, X, xsr, xsc, W, alpha, beta, Y, ysr, ysc);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_spd_inverse_f64(int32_t batch, int32_t k, const double* A, int64_t lda, int64_t stride, double jitter_rel,
                       double jitter_abs, double* linv, double* inv64, float* inv32, float* linvT32, int32_t* status,
                       void* stream) {
  if (batch < 1 || k < 1 || k > kSpdMax || !A || lda < k || stride < (int64_t)lda * (k - 1) + k || !linv || !status)
    return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(spd_inv_kernel, dim3(batch), dim3(256), 0, (hipStream_t)stream, k, A, lda, stride, jitter_rel,
                     jitter_abs, linv, inv64, inv32, linvT32, status);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_energy_loss_f64(int32_t k, int32_t m, const double* Sinv, const double* T, const double* w, float* C1,
                       float* C2, double* loss, void* stream) {
  if (k < 1 || k > kSpdMax || m < 1 || m > kSpdMax || !Sinv || !T || !w || !C1 || !C2 || !loss) return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(energy_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, k, m, Sinv, T, w, C1, C2, loss);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

size_t ck_pivchol_workspace_bytes(int64_t n) {
  if (n < 1) return 0;
  const size_t parts = (size_t)cdiv(n, kPcRows);
  return sizeof(float) * (size_t)n + parts * (sizeof(float) + sizeof(int)) + 16;
}

int ck_pivchol_dense_f32(int64_t n, const float* A, int64_t lda, int32_t rank, float* L, int64_t ldl, int32_t* piv,
                         float* dpiv, void* work, size_t work_bytes, void* stream) {
  if (n < 1 || n > 0x7FFFFFFF || !A || lda < n || rank < 1 || rank > n || !L || ldl < n || !piv || !dpiv)
    return CK_ERR_BAD_ARG;
  if (!work || work_bytes < ck_pivchol_workspace_bytes(n)) return CK_ERR_WORKSPACE;
  const int nparts = (int)cdiv(n, kPcRows);
  float* d = static_cast<float*>(work);
  float* pval = d + n;
  int* pidx = reinterpret_cast<int*>(pval + nparts);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(pc_init_kernel, dim3(nparts), dim3(256), 0, s, (int)n, A, lda, d, pval, pidx);
  for (int j = 0; j < rank; ++j) {
    hipLaunchKernelGGL(pc_select_kernel, dim3(1), dim3(64), 0, s, nparts, (const float*)pval, (const int*)pidx, j,
                       piv, dpiv);
    hipLaunchKernelGGL(pc_column_kernel, dim3(nparts), dim3(256), 0, s, (int)n, A, lda, j, (const int*)piv,
                       (const float*)dpiv, L, ldl, d, pval, pidx);
  }
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_rbf_dense_f32(int64_t n, int32_t dim, const float* P, int64_t ldp, float lengthscale, float* K, int64_t ldk,
                     uint16_t* Kb, int64_t ldkb, void* stream) {
  if (n < 1 || n > 0x7FFFFFFF || dim < 1 || dim > kMaxIn || !P || ldp < dim || !(lengthscale > 0.f) ||
      (!K && !Kb) || (K && ldk < n) || (Kb && ldkb < n))
    return CK_ERR_BAD_ARG;
  if (cdiv(n, 1024) > 65535) return CK_ERR_BAD_ARG;
  const float inv2l2 = 1.f / (2.f * lengthscale * lengthscale);
  hipLaunchKernelGGL(rbf_dense_kernel, dim3((unsigned)n, (unsigned)cdiv(n, 1024)), dim3(256), 0, (hipStream_t)stream,
                     (int)n, dim, P, ldp, inv2l2, K, ldk, Kb, ldkb);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_dense_apply_f64acc(int64_t n, const float* A, int64_t lda, double scale, double shift, int32_t nrhs,
                          const double* X, int64_t ldx, double* Y, int64_t ldy, void* stream) {
  if (n < 1 || n > 0x7FFFFFFF || !A || lda < n || nrhs < 1 || !X || !Y || ldx < n || ldy < n) return CK_ERR_BAD_ARG;
  hipStream_t s = (hipStream_t)stream;
  const unsigned grid = (unsigned)cdiv(n, 4 * kF64Rows);
  for (int b0 = 0; b0 < nrhs; b0 += kF64NB) {
    const int nb = nrhs - b0 < kF64NB ? nrhs - b0 : kF64NB;
    hipLaunchKernelGGL(dense_f64acc_kernel, dim3(grid), dim3(256), 0, s, (int)n, A, lda, X + (int64_t)b0 * ldx, ldx,
                       Y + (int64_t)b0 * ldy, ldy, nb, scale, shift);
  }
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // extern "C"
