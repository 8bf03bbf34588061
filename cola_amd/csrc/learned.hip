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

// xs[r][d] = X[(64 b + r) ldx + d] for the block's points, zero past n and din; the
// loads are unconditional (indexes clamped) so they issue together
__device__ __forceinline__ void load_points(float (*xs)[kMaxIn], const float* __restrict__ X, int64_t ldx, int n,
                                            int din, int t) {
#pragma unroll
  for (int k = 0; k < kRows * kMaxIn / 256; ++k) {
    const int e = t + 256 * k, rr = e / kMaxIn, d = e % kMaxIn;
    const int64_t i = (int64_t)blockIdx.x * kRows + rr;
    const float v = X[min(i, (int64_t)n - 1) * ldx + min(d, din - 1)];
    xs[rr][d] = (i < n && d < din) ? v : 0.f;
  }
}

// ------------------------------------------------------------------ MLP forward
// Weights of layers 2 and 3 are staged in LDS once per block and read back as
// wave-wide broadcasts; layer 1 (din <= 8 rows) reads W1 through scalar loads.
// Summation order per output is the input index order.
// One launch runs a few blocks once each, so its instruction cache starts cold: the
// activation epilogues are rolled loops over LDS and layer 3 runs in 8-column chunks,
// which keeps the code small (its size, not the arithmetic, set the time of an
// unrolled form).
template <bool SAVE>
__global__ __launch_bounds__(256) void mlp_fwd_kernel(int n, int din, int dout, const float* __restrict__ X,
                                                      int64_t ldx, const float* __restrict__ prm,
                                                      float* __restrict__ G, int64_t gsr, int64_t gsc,
                                                      float* __restrict__ Z1, float* __restrict__ Z2, int64_t ldz) {
  __shared__ float xs[kRows][kMaxIn];
  __shared__ float h1[kRows][kPitch];
  __shared__ float h2[kRows][kPitch];
  __shared__ __align__(16) float ws[kH * kMaxOut];  // W2 (kH x kH), then W3 (kH x dout), row-major
  const MlpLayout L = mlp_layout(din, dout);
  const int t = threadIdx.x;
  // g is the wave index: made wave-uniform (an SGPR), so the W1 loads below, whose
  // addresses depend only on g and the loop counters, become scalar loads
  const int r = t & 63, g = __builtin_amdgcn_readfirstlane(t >> 6);
  const int64_t row = (int64_t)blockIdx.x * kRows + r;
  const bool valid = row < n;
  // W2 and W3 start at multiples of kH floats (16-byte aligned); kH x dout is a
  // multiple of 4 floats
  constexpr int kQ = kH * kH / 4 / 256;  // float4 per thread for W2
  const float4* w2g = reinterpret_cast<const float4*>(prm + L.w2);
  const float4* w3g = reinterpret_cast<const float4*>(prm + L.w3);
  float4* ws4 = reinterpret_cast<float4*>(ws);
#pragma unroll
  for (int q = 0; q < kQ; ++q) ws4[t + 256 * q] = w2g[t + 256 * q];
  load_points(xs, X, ldx, n, din, t);
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
    for (int cc = 0; cc < 32; ++cc) h1[r][c0 + cc] = acc[cc];
  }
  // each lane activates its own row's 32 entries: no barrier needed
#pragma unroll 1
  for (int cc = 0; cc < 32; ++cc) {
    const float z = h1[r][c0 + cc];
    if (SAVE && valid) Z1[(int64_t)(c0 + cc) * ldz + row] = z;
    h1[r][c0 + cc] = gelu(z);
  }
  __syncthreads();
  {
    float acc[32];
#pragma unroll
    for (int cc = 0; cc < 32; ++cc) acc[cc] = prm[L.b2 + c0 + cc];
    for (int j = 0; j < kH; ++j) {
      const float hv = h1[r][j];
      const float4* w = reinterpret_cast<const float4*>(ws + j * kH + c0);
#pragma unroll
      for (int c4 = 0; c4 < 8; ++c4) {
        const float4 wv = w[c4];
        acc[4 * c4 + 0] = fmaf(hv, wv.x, acc[4 * c4 + 0]);
        acc[4 * c4 + 1] = fmaf(hv, wv.y, acc[4 * c4 + 1]);
        acc[4 * c4 + 2] = fmaf(hv, wv.z, acc[4 * c4 + 2]);
        acc[4 * c4 + 3] = fmaf(hv, wv.w, acc[4 * c4 + 3]);
      }
    }
#pragma unroll
    for (int cc = 0; cc < 32; ++cc) h2[r][c0 + cc] = acc[cc];
  }
#pragma unroll 1
  for (int cc = 0; cc < 32; ++cc) {
    const float z = h2[r][c0 + cc];
    if (SAVE && valid) Z2[(int64_t)(c0 + cc) * ldz + row] = z;
    h2[r][c0 + cc] = gelu(z);
  }
  __syncthreads();  // W2 reads done, h2 complete
  const int n3 = kH * dout / 4;  // float4 of W3, <= 256 kQ
#pragma unroll
  for (int q = 0; q < kQ; ++q)
    if (t + 256 * q < n3) ws4[t + 256 * q] = w3g[t + 256 * q];
  __syncthreads();
  // layer 3: wave g owns outputs o0 .. o0 + per - 1, in chunks of 8. A chunk's tail
  // past its owned range reads other LDS words (within ws: o0 + c8 + 7 <= 4 per + 6
  // and row 127 ends at 128 dout) into accumulators that are never stored.
  const int per = (dout + 3) / 4;  // <= 32 outputs per wave
  const int o0 = g * per;
#pragma unroll 1
  for (int c8 = 0; c8 < per; c8 += 8) {
    float acc[8];
#pragma unroll
    for (int cc = 0; cc < 8; ++cc)
      acc[cc] = (c8 + cc < per && o0 + c8 + cc < dout) ? prm[L.b3 + o0 + c8 + cc] : 0.f;
    for (int j = 0; j < kH; ++j) {
      const float hv = h2[r][j];
      const float* w = ws + j * dout + o0 + c8;
#pragma unroll
      for (int cc = 0; cc < 8; ++cc) acc[cc] = fmaf(hv, w[cc], acc[cc]);
    }
    if (valid) {
#pragma unroll
      for (int cc = 0; cc < 8; ++cc)
        if (c8 + cc < per && o0 + c8 + cc < dout) G[row * gsr + (int64_t)(o0 + c8 + cc) * gsc] = acc[cc];
    }
  }
}

// ----------------------------------------------------------------- MLP backward
// Given dL/dG for a block of 64 points, writes that block's contribution to every
// parameter gradient (flat layout) into part[blockIdx.x * total + ...].
// W3 (row pitch dout rounded up to 4, zero-padded) and then W2 are staged in LDS;
// dZ2 and dZ1 are accumulated in registers over 4 columns at a time, with the weight
// rows read as wave-wide float4 broadcasts. dG and dZ2 share one LDS array (dG is
// dead once dZ2 is in registers). The saved pre-activations are loaded into
// registers up front (one latency); GELU and its derivative run in rolled loops over
// LDS so the code stays small. Summation order per output is the index order.
__global__ __launch_bounds__(256) void mlp_bwd_kernel(int n, int din, int dout, const float* __restrict__ X,
                                                      int64_t ldx, const float* __restrict__ prm,
                                                      const float* __restrict__ Z1, const float* __restrict__ Z2,
                                                      int64_t ldz, const float* __restrict__ dG, int64_t dsr,
                                                      int64_t dsc, float* __restrict__ part) {
  __shared__ float xs[kRows][kMaxIn];
  __shared__ float a[kRows][kPitch];   // Z2 -> h2; Z2; Z1 -> h1; dZ1
  __shared__ float dd[kRows][kPitch];  // dG (columns up to dp, zero beyond dout); dZ2; Z1
  __shared__ __align__(16) float ws[kH * kMaxOut];  // W3 (kH x dp), then W2 (kH x kH)
  const MlpLayout L = mlp_layout(din, dout);
  float* out = part + (int64_t)blockIdx.x * L.total;
  const int t = threadIdx.x;
  const int r = t & 63, g = __builtin_amdgcn_readfirstlane(t >> 6);
  const int64_t row = (int64_t)blockIdx.x * kRows + r;
  const bool valid = row < n;
  const int dp = (dout + 3) & ~3;
  const int c0 = g * 32;  // this wave's 32 hidden units
  constexpr int kQ = kH * kH / 4 / 256;  // float4 of W2 (and at most of W3) per thread
  // every global load below is unconditional (row and column indexes clamped into
  // range, the value selected afterwards), so each group is issued back to back and
  // waited for once
  const int64_t rowc = valid ? row : (int64_t)n - 1;
  float z2[32], z1[32], gv[32];
#pragma unroll
  for (int jj = 0; jj < 32; ++jj) {
    const float v2 = Z2[(int64_t)(c0 + jj) * ldz + rowc], v1 = Z1[(int64_t)(c0 + jj) * ldz + rowc];
    z2[jj] = valid ? v2 : 0.f;
    z1[jj] = valid ? v1 : 0.f;
  }
#pragma unroll
  for (int i = 0; i < 32; ++i) {  // dG columns c = g + 4 i
    const int c = g + 4 * i;
    const float v = dG[rowc * dsr + (int64_t)min(c, dout - 1) * dsc];
    gv[i] = (valid && c < dout) ? v : 0.f;
  }
  load_points(xs, X, ldx, n, din, t);
  if (dp == dout) {  // W3 rows are 16-byte aligned in both layouts: straight float4 copy
    const float4* w3g = reinterpret_cast<const float4*>(prm + L.w3);
    float4* ws4 = reinterpret_cast<float4*>(ws);
    const int n3 = kH * dout / 4;
#pragma unroll
    for (int q = 0; q < kQ; ++q)
      if (t + 256 * q < n3) ws4[t + 256 * q] = w3g[t + 256 * q];
  } else {
    // lane r copies columns r and r + 64 of rows g, g + 4, ...; zero padding to dp
#pragma unroll 1
    for (int j0 = 0; j0 < kH; j0 += 32) {
      float v0[8], v1[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const float* w = prm + L.w3 + (int64_t)(j0 + 4 * i + g) * dout;
        const float u0 = w[min(r, dout - 1)], u1 = w[min(r + 64, dout - 1)];
        v0[i] = r < dout ? u0 : 0.f;
        v1[i] = r + 64 < dout ? u1 : 0.f;
      }
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        float* w = ws + (j0 + 4 * i + g) * dp;
        if (r < dp) w[r] = v0[i];
        if (r + 64 < dp) w[r + 64] = v1[i];
      }
    }
  }
  // columns g + 4 i reach 127 < kPitch; those past dp receive zeros and are never read
  // before dd is rewritten, so the stores need no test
#pragma unroll
  for (int i = 0; i < 32; ++i) dd[r][g + 4 * i] = gv[i];
#pragma unroll
  for (int jj = 0; jj < 32; ++jj) a[r][c0 + jj] = z2[jj];
#pragma unroll 1
  for (int jj = 0; jj < 32; ++jj) a[r][c0 + jj] = gelu(a[r][c0 + jj]);  // own row: no barrier
  __syncthreads();

  // dW3[j][c] = sum_r h2[r][j] dG[r][c]; db3[c] = sum_r dG[r][c]
  {
    const int j = t & 127, par = t >> 7;
    const int nq = (dout - par + 1) / 2;  // columns c = par + 2 q
#pragma unroll 1
    for (int q0 = 0; q0 < nq; q0 += 16) {
      float acc[16];
#pragma unroll
      for (int qq = 0; qq < 16; ++qq) acc[qq] = 0.f;
      // columns past dout are clamped to the pad column kH; those sums are not stored
      for (int rr = 0; rr < kRows; ++rr) {
        const float av = a[rr][j];
#pragma unroll
        for (int qq = 0; qq < 16; ++qq) acc[qq] = fmaf(av, dd[rr][min(par + 2 * (q0 + qq), kH)], acc[qq]);
      }
#pragma unroll
      for (int qq = 0; qq < 16; ++qq)
        if (q0 + qq < nq) out[L.w3 + (int64_t)j * dout + par + 2 * (q0 + qq)] = acc[qq];
    }
    for (int c = t; c < dout; c += 256) {
      float s = 0.f;
      for (int rr = 0; rr < kRows; ++rr) s += dd[rr][c];
      out[L.b3 + c] = s;
    }
  }
  // dZ2[r][j] = (sum_c dG[r][c] W3[j][c]) * gelu'(Z2[r][j]), j = c0 + jj
  float acc[32];
#pragma unroll
  for (int jj = 0; jj < 32; ++jj) acc[jj] = 0.f;
  for (int c = 0; c < dp; c += 4) {
    const float x0 = dd[r][c], x1 = dd[r][c + 1], x2 = dd[r][c + 2], x3 = dd[r][c + 3];
#pragma unroll
    for (int jj = 0; jj < 32; ++jj) {
      const float4 w = *reinterpret_cast<const float4*>(ws + (c0 + jj) * dp + c);
      acc[jj] = fmaf(x0, w.x, acc[jj]);
      acc[jj] = fmaf(x1, w.y, acc[jj]);
      acc[jj] = fmaf(x2, w.z, acc[jj]);
      acc[jj] = fmaf(x3, w.w, acc[jj]);
    }
  }
  __syncthreads();  // dG, h2 and W3 reads done
  {  // W2 starts 16-byte aligned
    const float4* w2g = reinterpret_cast<const float4*>(prm + L.w2);
    float4* ws4 = reinterpret_cast<float4*>(ws);
#pragma unroll
    for (int q = 0; q < kQ; ++q) ws4[t + 256 * q] = w2g[t + 256 * q];
  }
#pragma unroll
  for (int jj = 0; jj < 32; ++jj) {
    dd[r][c0 + jj] = acc[jj];
    a[r][c0 + jj] = z2[jj];
  }
#pragma unroll 1
  for (int jj = 0; jj < 32; ++jj) {
    const float v = dd[r][c0 + jj];
    dd[r][c0 + jj] = valid ? v * gelu_grad(a[r][c0 + jj]) : 0.f;
  }
#pragma unroll
  for (int jj = 0; jj < 32; ++jj) a[r][c0 + jj] = z1[jj];
#pragma unroll 1
  for (int jj = 0; jj < 32; ++jj) a[r][c0 + jj] = gelu(a[r][c0 + jj]);
  __syncthreads();
  // dW2[j][k] = sum_r h1[r][j] dZ2[r][k]; db2[k] = sum_r dZ2[r][k]
  {
    const int j = t & 127, par = t >> 7;
#pragma unroll 1
    for (int q0 = 0; q0 < kH / 2; q0 += 16) {
      float acc2[16];
#pragma unroll
      for (int qq = 0; qq < 16; ++qq) acc2[qq] = 0.f;
      for (int rr = 0; rr < kRows; ++rr) {
        const float av = a[rr][j];
#pragma unroll
        for (int qq = 0; qq < 16; ++qq) acc2[qq] = fmaf(av, dd[rr][par + 2 * (q0 + qq)], acc2[qq]);
      }
#pragma unroll
      for (int qq = 0; qq < 16; ++qq) out[L.w2 + (int64_t)j * kH + par + 2 * (q0 + qq)] = acc2[qq];
    }
    for (int k = t; k < kH; k += 256) {
      float s = 0.f;
      for (int rr = 0; rr < kRows; ++rr) s += dd[rr][k];
      out[L.b2 + k] = s;
    }
  }
  // dZ1[r][j] = (sum_k dZ2[r][k] W2[j][k]) * gelu'(Z1[r][j]), j = c0 + jj
#pragma unroll
  for (int jj = 0; jj < 32; ++jj) acc[jj] = 0.f;
  for (int k = 0; k < kH; k += 4) {
    const float x0 = dd[r][k], x1 = dd[r][k + 1], x2 = dd[r][k + 2], x3 = dd[r][k + 3];
#pragma unroll
    for (int jj = 0; jj < 32; ++jj) {
      const float4 w = *reinterpret_cast<const float4*>(ws + (c0 + jj) * kH + k);
      acc[jj] = fmaf(x0, w.x, acc[jj]);
      acc[jj] = fmaf(x1, w.y, acc[jj]);
      acc[jj] = fmaf(x2, w.z, acc[jj]);
      acc[jj] = fmaf(x3, w.w, acc[jj]);
    }
  }
  __syncthreads();  // h1 and dZ2 reads done
#pragma unroll
  for (int jj = 0; jj < 32; ++jj) {
    a[r][c0 + jj] = acc[jj];
    dd[r][c0 + jj] = z1[jj];
  }
#pragma unroll 1
  for (int jj = 0; jj < 32; ++jj) {
    const float v = a[r][c0 + jj];
    a[r][c0 + jj] = valid ? v * gelu_grad(dd[r][c0 + jj]) : 0.f;
  }
  __syncthreads();
  // dW1[d][j] = sum_r x[r][d] dZ1[r][j]; db1[j] = sum_r dZ1[r][j]
  {
    const int j = t & 127, par = t >> 7;
    for (int d = par; d < din; d += 2) {
      float s = 0.f;
      for (int rr = 0; rr < kRows; ++rr) s = fmaf(xs[rr][d], a[rr][j], s);
      out[L.w1 + (int64_t)d * kH + j] = s;
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
constexpr int kGramChunk = 64;        // rows staged per pass
constexpr int64_t kGramMaxParts = 1024;  // partials per output entry, at most
// rows per partial: one 64-row chunk per block until n needs more than kGramMaxParts
inline int64_t gram_rows(int64_t n) { return kGramChunk * cdiv(cdiv(n, kGramChunk), kGramMaxParts); }

__global__ __launch_bounds__(256) void gram_partials(int64_t n, int64_t rows, int ka, int kb,
                                                     const float* __restrict__ X, int64_t xsr, int64_t xsc,
                                                     const float* __restrict__ Y, int64_t ysr, int64_t ysc,
                                                     double* __restrict__ part) {
  __shared__ float xs[64][33];
  __shared__ float ys[64][33];
  const int ta = blockIdx.y, tb = blockIdx.z;
  const int al = threadIdx.x >> 3, bq = (threadIdx.x & 7) * 4;
  const int64_t r0 = (int64_t)blockIdx.x * rows;
  const int64_t r1 = r0 + rows < n ? r0 + rows : n;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t rb = r0; rb < r1; rb += kGramChunk) {
    // loads unconditional (indexes clamped) so a chunk's loads issue together
    float vx[8], vy[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = threadIdx.x + 256 * q, rr = e & 63, cc = e >> 6;
      const int64_t i = rb + rr < r1 ? rb + rr : r1 - 1;
      vx[q] = X[i * xsr + (int64_t)min(ta * 32 + cc, ka - 1) * xsc];
      vy[q] = Y[i * ysr + (int64_t)min(tb * 32 + cc, kb - 1) * ysc];
    }
    __syncthreads();  // previous chunk's reads done
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = threadIdx.x + 256 * q, rr = e & 63, cc = e >> 6;
      const bool in = rb + rr < r1;
      xs[rr][cc] = (in && ta * 32 + cc < ka) ? vx[q] : 0.f;
      ys[rr][cc] = (in && tb * 32 + cc < kb) ? vy[q] : 0.f;
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
  // W's column slab in groups of eight loads per thread (indexes clamped, so a group
  // issues together)
  for (int e0 = 0; e0 < kin * 32; e0 += 2048) {
    float v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = min(e0 + (int)threadIdx.x + 256 * q, kin * 32 - 1), j = e / 32, cc = e % 32;
      v[q] = W[(int64_t)j * kout + c0 + min(cc, nc - 1)];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = e0 + threadIdx.x + 256 * q;
      if (e < kin * 32) ws[e] = e % 32 < nc ? v[q] : 0.f;
    }
  }
  __syncthreads();
  if (i >= n) return;
  float acc[32];
#pragma unroll
  for (int cc = 0; cc < 32; ++cc) acc[cc] = 0.f;
  // X row entries eight at a time so their loads are in flight together
  int j = 0;
  for (; j + 8 <= kin; j += 8) {
    float xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) xv[u] = X[i * xsr + (int64_t)(j + u) * xsc];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int cc = 0; cc < 32; ++cc) acc[cc] = fmaf(xv[u], ws[(j + u) * 32 + cc], acc[cc]);
  }
  for (; j < kin; ++j) {
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
  __shared__ double dgl[kSpdMax];  // Cholesky diagonal, kept apart until the factorization ends
  __shared__ double shift;
  const int b = blockIdx.x, t = threadIdx.x;
  const int kk = k * k;
  const double* Ab = A + (int64_t)b * sa;
  // lower triangle in; loads are unconditional (index clamped) so each group of eight
  // issues together
  for (int e0 = 0; e0 < kk; e0 += 2048) {
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = min(e0 + t + 256 * q, kk - 1), i = e / k, j = e % k;
      v[q] = Ab[(int64_t)i * lda + j];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = e0 + t + 256 * q, i = e / k, j = e % k;
      if (e < kk && j <= i) Ls[pk(i, j)] = v[q];
    }
  }
  for (int e = t; e < k * (k + 1) / 2; e += 256) Xs[e] = 0.0;
  __syncthreads();
  if (t == 0) {
    double tr = 0.0;
    for (int i = 0; i < k; ++i) tr += Ls[pk(i, i)];
    shift = jitter_abs + jitter_rel * tr / k;
  }
  __syncthreads();
  if (t < k) Ls[pk(t, t)] += shift;
  // this thread's entries of the packed lower triangle, e = t + 256 m, as (row, column);
  // the per-step updates below touch only these, unrolled so their LDS accesses issue
  // back to back
  constexpr int kM = (kPack + 255) / 256;
  const int np = k * (k + 1) / 2;
  int er[kM], ec[kM], rb[kM], cb[kM];  // row, column, and the packed offsets of their rows
#pragma unroll
  for (int m = 0; m < kM; ++m) {
    const int e = t + 256 * m;
    int i = (int)((sqrtf(8.f * (float)e + 1.f) - 1.f) * 0.5f);
    while (i * (i + 1) / 2 > e) --i;
    while ((i + 1) * (i + 2) / 2 <= e) ++i;
    er[m] = e < np ? i : -1;
    ec[m] = e < np ? e - i * (i + 1) / 2 : kSpdMax;
    rb[m] = e < np ? i * (i + 1) / 2 : 0;
    cb[m] = e < np ? ec[m] * (ec[m] + 1) / 2 : 0;
  }
  // right-looking Cholesky, two barriers per column: every thread reads the pivot
  // and takes its square root itself (the same value everywhere, so the break on a
  // non-positive pivot is block-uniform); thread 0 records it in dgl
  int bad = 0;
  for (int j = 0; j < k; ++j) {
    __syncthreads();
    const double v = Ls[pk(j, j)];
    if (!(v > 0.0)) {
      bad = j + 1;
      break;
    }
    const double djj = sqrt(v);
    if (t == 0) dgl[j] = djj;
    for (int i = j + 1 + t; i < k; i += 256) Ls[pk(i, j)] /= djj;
    __syncthreads();
    // trailing update of the lower triangle: entries (i, l) with j < l <= i. It reads
    // column j and writes only columns > j, so all reads go first (unconditional: the
    // offsets stay inside the triangle) and the writes after
    double lv[kM], rv[kM], ov[kM];
#pragma unroll
    for (int m = 0; m < kM; ++m) {
      lv[m] = Ls[rb[m] + min(j, er[m] < 0 ? 0 : er[m])];
      rv[m] = Ls[cb[m] + (ec[m] < kSpdMax ? min(j, ec[m]) : 0)];
      ov[m] = Ls[min(t + 256 * m, kPack - 1)];
    }
#pragma unroll
    for (int m = 0; m < kM; ++m)
      if (ec[m] > j && ec[m] < kSpdMax) Ls[t + 256 * m] = ov[m] - lv[m] * rv[m];
  }
  if (t == 0) status[b] = bad;
  if (bad) return;
  __syncthreads();
  if (t < k) Ls[pk(t, t)] = dgl[t];
  // X = L^-1, one row per step: row q is final once the steps before it have added
  // their terms, X[q][c] = -(sum_{c<=p<q} L[q][p] X[p][c]) / L[q][q]; the step then adds
  // L[i][q] X[q][c] to every later row i. Each sum takes its terms in increasing p,
  // the order of a column-by-column forward substitution.
  for (int q = 0; q < k; ++q) {
    __syncthreads();
    if (t <= q) Xs[pk(q, t)] = t == q ? 1.0 / Ls[pk(q, q)] : -Xs[pk(q, t)] / Ls[pk(q, q)];
    __syncthreads();
    // rows > q are written, row q and L are read: reads first, then writes
    const int qb = q * (q + 1) / 2;
    double lv[kM], xv[kM], ov[kM];
#pragma unroll
    for (int m = 0; m < kM; ++m) {
      lv[m] = Ls[rb[m] + min(q, er[m] < 0 ? 0 : er[m])];
      xv[m] = Xs[qb + min(ec[m], q)];
      ov[m] = Xs[min(t + 256 * m, kPack - 1)];
    }
#pragma unroll
    for (int m = 0; m < kM; ++m)
      if (er[m] > q && ec[m] <= q) Xs[t + 256 * m] = fma(lv[m], xv[m], ov[m]);
  }
  __syncthreads();
  double* Li = linv + (int64_t)b * kk;
  for (int e = t; e < kk; e += 256) {
    const int r = e / k, cc = e % k;
    Li[e] = cc <= r ? Xs[pk(r, cc)] : 0.0;
    if (linvT32) linvT32[(int64_t)b * kk + e] = r <= cc ? (float)Xs[pk(cc, r)] : 0.f;
  }
  if (inv64 || inv32) {
    // S^-1 = X^T X: thread owns the 4 x 4 block at rows r0.., columns c0..; entry (r, c)
    // sums X[q][r] X[q][c] over q >= max(r, c) in increasing q (the block's earlier q
    // add exact zeros, as X[q][r] = 0 for r > q)
    const int r0 = 4 * (t >> 4), c0 = 4 * (t & 15);
    if (r0 < k && c0 < k) {
      double acc[4][4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int w = 0; w < 4; ++w) acc[u][w] = 0.0;
#pragma unroll 2
      for (int q = r0 > c0 ? r0 : c0; q < k; ++q) {
        double xr[4], xc[4];
        const int base = q * (q + 1) / 2;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          xr[u] = r0 + u <= q ? Xs[base + r0 + u] : 0.0;
          xc[u] = c0 + u <= q ? Xs[base + c0 + u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int w = 0; w < 4; ++w) acc[u][w] = fma(xr[u], xc[w], acc[u][w]);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int w = 0; w < 4; ++w) {
          if (r0 + u >= k || c0 + w >= k) continue;
          const int64_t e = (int64_t)b * kk + (r0 + u) * k + c0 + w;
          if (inv64) inv64[e] = acc[u][w];
          if (inv32) inv32[e] = (float)acc[u][w];
        }
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
  __shared__ double Ss[kSpdMax * kSpdMax];  // S^-1 rebuilt symmetric from its lower triangle
  __shared__ double ws[kSpdMax];
  __shared__ double red[4];
  const int t = threadIdx.x;
  // loads in groups of eight per thread, indexes clamped so each group issues together
  for (int e0 = 0; e0 < kSpdMax * kSpdMax; e0 += 2048) {
    double vt[8], vs[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = e0 + t + 256 * q;
      const int et = min(e, m * k - 1), es = min(e, k * k - 1), i = es / k, j = es % k;
      vt[q] = T[et];
      vs[q] = Sinv[i >= j ? es : j * k + i];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = e0 + t + 256 * q;
      if (e < m * k) Ts[e] = vt[q];
      if (e < k * k) Ss[e] = vs[q];
    }
  }
  if (t < m) ws[t] = w[t];
  __syncthreads();
  double wsum = 0.0;
  for (int i = 0; i < m; ++i) wsum += ws[i];
  const double c = 2.0 / wsum;
  // P[i][a] = sum_q T[i][q] S^-1[q][a] for a = g + 4 u: sixteen independent sums per
  // thread, each over q in increasing order
  const int i = t >> 2, g = t & 3;
  double pv[kSpdMax / 4];
  double f = 0.0;
  if (i < m) {
#pragma unroll
    for (int u = 0; u < kSpdMax / 4; ++u) pv[u] = 0.0;
    for (int q = 0; q < k; ++q) {
      const double tq = Ts[i * k + q];
      const double* srow = Ss + q * k;
#pragma unroll
      for (int u = 0; u < kSpdMax / 4; ++u)
        if (g + 4 * u < k) pv[u] = fma(tq, srow[g + 4 * u], pv[u]);
    }
#pragma unroll
    for (int u = 0; u < kSpdMax / 4; ++u) {
      const int a = g + 4 * u;
      if (a < k) {
        C1[i * k + a] = (float)(-c * ws[i] * pv[u]);
        f = fma(ws[i] * pv[u], Ts[i * k + a], f);
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
  // C2[a][b] = c sum_r (w_r P[r][a]) P[r][b]: thread owns the 4 x 4 block at (a0, b0)
  const int a0 = 4 * (t >> 4), b0 = 4 * (t & 15);
  if (a0 < k && b0 < k) {
    double acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
    for (int r = 0; r < m; ++r) {
      const double* prow = Ts + r * k;
      double wa[4], pb[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        wa[u] = ws[r] * prow[min(a0 + u, k - 1)];
        pb[u] = prow[min(b0 + u, k - 1)];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = fma(wa[u], pb[v], acc[u][v]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (a0 + u < k && b0 + v < k) C2[(a0 + u) * k + b0 + v] = (float)(c * acc[u][v]);
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

// The same product with every load unconditional (rows and columns clamped, the extra terms
// multiplied by zero) and the x chunk staged once per block for all of its rows: eight waves of
// four rows share one FP64 copy of X[:, c0:c0+512) in LDS; A is streamed with non-temporal
// float4 loads.  Requires n % 4 == 0, 16-byte aligned rows of A and X, lda % 4 == 0 and
// ldx % 2 == 0 (the entry point keeps dense_f64acc_kernel for the rest).
constexpr int kF64Chunk = 512;
constexpr int kF64Waves = 8;
typedef float f64acc_f4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(64 * kF64Waves) __attribute__((amdgpu_waves_per_eu(4))) void dense_f64acc_lds_kernel(int n, const float* __restrict__ A,
                                                                         int64_t lda, const double* __restrict__ X,
                                                                         int64_t ldx, double* __restrict__ Y,
                                                                         int64_t ldy, int nb, double scale,
                                                                         double shift) {
  __shared__ double2 xs[kF64NB][kF64Chunk / 2];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // rows and their pointers stay scalar
  const int row0 = blockIdx.x * (4 * kF64Waves) + wave * 4;
  double acc[4][kF64NB];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int b = 0; b < kF64NB; ++b) acc[r][b] = 0.0;
  const float* arow[4];
  float rmask[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = row0 + r < n ? row0 + r : n - 1;
    arow[r] = A + (int64_t)row * lda;
    rmask[r] = row0 + r < n ? 1.f : 0.f;
  }
  for (int c0 = 0; c0 < n; c0 += kF64Chunk) {
    const int cn = n - c0 < kF64Chunk ? n - c0 : kF64Chunk;  // a multiple of 4
    __syncthreads();
    for (int e = threadIdx.x; e < kF64NB * (kF64Chunk / 2); e += 64 * kF64Waves) {
      const int b = e / (kF64Chunk / 2), j2 = e % (kF64Chunk / 2);
      xs[b][j2] = (b < nb && 2 * j2 < cn) ? *reinterpret_cast<const double2*>(X + (int64_t)b * ldx + c0 + 2 * j2)
                                          : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kF64Chunk / 256; ++q) {
      const int j = q * 256 + lane * 4;
      const int jc = j < cn ? j : cn - 4;  // clamped: the load stays unconditional
      float4 a[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const f64acc_f4 v = __builtin_nontemporal_load(reinterpret_cast<const f64acc_f4*>(arow[r] + c0 + jc));
        const float m = j < cn ? rmask[r] : 0.f;
        a[r] = make_float4(v.x * m, v.y * m, v.z * m, v.w * m);
      }
#pragma unroll
      for (int b = 0; b < kF64NB; ++b) {
        const double2 x0 = xs[b][jc / 2], x1 = xs[b][jc / 2 + 1];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          double t = acc[r][b];
          t = fma((double)a[r].x, x0.x, t);
          t = fma((double)a[r].y, x0.y, t);
          t = fma((double)a[r].z, x1.x, t);
          t = fma((double)a[r].w, x1.y, t);
          acc[r][b] = t;
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
#pragma unroll
    for (int b = 0; b < kF64NB; ++b) {
      const double sum = wave_sum(acc[r][b]);
      const int row = row0 + r;
      if (lane == 0 && row < n && b < nb)
        Y[(int64_t)b * ldy + row] = scale * sum + shift * X[(int64_t)b * ldx + row];
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
  if (reinterpret_cast<uintptr_t>(params) % 16) return CK_ERR_BAD_ARG;  // weights are staged as float4
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
  if (reinterpret_cast<uintptr_t>(params) % 16) return CK_ERR_BAD_ARG;  // weights are staged as float4
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
  return sizeof(double) * (size_t)cdiv(n, gram_rows(n)) * (size_t)ka * (size_t)kb;
}

int ck_gram_f64(int64_t n, int32_t ka, int32_t kb, const float* X, int64_t xsr, int64_t xsc, const float* Y,
                int64_t ysr, int64_t ysc, double* C, void* work, size_t work_bytes, void* stream) {
  if (n < 1 || ka < 1 || kb < 1 || !X || !Y || !C) return CK_ERR_BAD_ARG;
  if (!work || work_bytes < ck_gram_workspace_bytes(n, ka, kb)) return CK_ERR_WORKSPACE;
  const int64_t rows = gram_rows(n);
  const int nch = (int)cdiv(n, rows);
  hipStream_t s = (hipStream_t)stream;
  double* part = static_cast<double*>(work);
  hipLaunchKernelGGL(gram_partials, dim3(nch, (unsigned)cdiv(ka, 32), (unsigned)cdiv(kb, 32)), dim3(256), 0, s, n,
                     rows, ka, kb, X, xsr, xsc, Y, ysr, ysc, part);
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
  auto al16 = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
  const bool vec = n % 4 == 0 && lda % 4 == 0 && ldx % 2 == 0 && al16(A) && al16(X);
  const unsigned grid = (unsigned)cdiv(n, 4 * kF64Rows);
  const unsigned grid_lds = (unsigned)cdiv(n, 4 * kF64Waves);
  for (int b0 = 0; b0 < nrhs; b0 += kF64NB) {
    const int nb = nrhs - b0 < kF64NB ? nrhs - b0 : kF64NB;
    if (vec)
      hipLaunchKernelGGL(dense_f64acc_lds_kernel, dim3(grid_lds), dim3(64 * kF64Waves), 0, s, (int)n, A, lda,
                         X + (int64_t)b0 * ldx, ldx, Y + (int64_t)b0 * ldy, ldy, nb, scale, shift);
    else
      hipLaunchKernelGGL(dense_f64acc_kernel, dim3(grid), dim3(256), 0, s, (int)n, A, lda, X + (int64_t)b0 * ldx,
                         ldx, Y + (int64_t)b0 * ldy, ldy, nb, scale, shift);
  }
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // extern "C"
