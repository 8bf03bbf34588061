// Batched Lanczos (double classical Gram-Schmidt) and Arnoldi (modified or
// double classical Gram-Schmidt) on basis blocks V[b][j][i].
//
// Re-orthogonalisation is a batched GEMV pair per pass: h = V^T w over the filled
// columns only (the reference sweeps all m+2 preallocated columns, zeros
// included; lanczos.py:287-296), then w -= V h.  The projection kernel stages
// one 1024-element tile of w in LDS and lets each wavefront dot a subset of
// basis vectors against it, writing one partial per (column, tile); a small
// kernel adds the partials in index order.  All loop control (i, done, the
// beta-relative stop rule of lanczos.py:263-268) stays on the device.
#include "common.hpp"
#include "ops_internal.hpp"

using namespace ck;

namespace {

constexpr int kTile = 1024;  // vector elements per block in the sweeps

struct LzScratch {
  float* dpart;  // nb x P
  float* hpart;  // nb x (m+2) x P
  float* h;      // nb x (m+2)
  float* h2;     // nb x (m+2)   (second pass; Arnoldi keeps both)
  float* nrm;    // nb
  int P;
  int M2;        // m + 2
};

inline int lz_nparts(int64_t n) { return (int)cdiv(n, kTile); }

inline LzScratch scratch_of(void* scratch, int64_t n, int nb, int m) {
  LzScratch s;
  s.P = lz_nparts(n);
  s.M2 = m + 2;
  float* f = static_cast<float*>(scratch);
  s.dpart = f;
  f += (size_t)nb * s.P * 2;  // two buffers (Arnoldi MGS alternates them)
  s.hpart = f;
  f += (size_t)nb * s.M2 * s.P;
  s.h = f;
  f += (size_t)nb * s.M2;
  s.h2 = f;
  f += (size_t)nb * s.M2;
  s.nrm = f;
  return s;
}

__device__ inline double col_total(const float* __restrict__ part, int P, int64_t base) {
  // each lane adds its partials lane, lane + 64, ... in that order; loads go four at a time
  double acc = 0.0;
  int i = threadIdx.x & 63;
  for (; i + 192 < P; i += 256) {
    const float v0 = part[base + i], v1 = part[base + i + 64], v2 = part[base + i + 128], v3 = part[base + i + 192];
    acc += (double)v0;
    acc += (double)v1;
    acc += (double)v2;
    acc += (double)v3;
  }
  for (; i < P; i += 64) acc += (double)part[base + i];
  return wave_sum(acc);
}

// sum of squares (MODE 1) or dot with Y (MODE 0) of one tile per block
template <int MODE>
__global__ __launch_bounds__(256) void tile_partials(int64_t n, const float* __restrict__ X, int64_t ldb_x,
                                                     const float* __restrict__ Y, int64_t ldb_y,
                                                     float* __restrict__ part, int P, const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ float red[4];
  const int b = blockIdx.y, blk = blockIdx.x;
  const int64_t i0 = (int64_t)blk * kTile;
  const float* x = X + b * ldb_x;
  const float* y = Y + b * ldb_y;
  float acc = 0.f;
  for (int e = threadIdx.x; e < kTile; e += 256) {
    const int64_t i = i0 + e;
    if (i < n) {
      const float a = x[i];
      acc = fmaf(a, MODE == 0 ? y[i] : a, acc);
    }
  }
  acc = block_sum<256>(acc, red);
  if (threadIdx.x == 0) part[(int64_t)b * P + blk] = acc;
}

// ---------------------------------------------------------------- Lanczos start
__global__ __launch_bounds__(256) void lz_norm_final(int nb, const float* __restrict__ part, int P,
                                                     float* __restrict__ nrm, const int* __restrict__ done) {
  if (done && *done) return;
  const int wave = threadIdx.x >> 6;
  for (int b = wave; b < nb; b += 4) {
    const double s = col_total(part, P, (int64_t)b * P);
    if ((threadIdx.x & 63) == 0) nrm[b] = (float)sqrt(s);
  }
}

// dst = src / nrm[b] (column 1 of V at start), and column 0 = 0
__global__ __launch_bounds__(256) void lz_start_fill(int64_t n, const float* __restrict__ s, int64_t lds_,
                                                     float* __restrict__ V, int64_t ldv, int64_t ldb,
                                                     const float* __restrict__ nrm) {
  const int b = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V[b * ldb + i] = 0.f;
  V[b * ldb + ldv + i] = s[b * lds_ + i] / nrm[b];
}

__global__ __launch_bounds__(256) void lz_start_ctrl(int nb, int m, float* __restrict__ diag,
                                                     float* __restrict__ subdiag, void* state) {
  ck_lz_ctrl* c = static_cast<ck_lz_ctrl*>(state);
  float* errors = reinterpret_cast<float*>(c + 1);
  for (int t = threadIdx.x; t < nb * m; t += 256) diag[t] = 0.f;
  for (int t = threadIdx.x; t < nb * (m + 1); t += 256) subdiag[t] = 0.f;
  if (threadIdx.x == 0) {
    c->i = 1;
    c->max_iters = m;
    c->nb = nb;
    c->done = (1 <= m) ? 0 : 1;  // lanczos.py:263-268 at i = 1
    errors[0] = 1.0f;            // lanczos.py:256-261: 0/max(0,1e-30) + (i<=1)
  }
}

// V[:,i] /= denom, denom = nrm[b] (first step) or subdiag[b][i-1]
// V4: four consecutive elements per thread as one 16-byte access (column start and probe
// stride multiples of 4 floats, V 16-byte aligned); the tail past n is left alone
template <bool V4>
__global__ __launch_bounds__(256) void lz_scale_col(int64_t n, float* __restrict__ V, int64_t ldb, int64_t coloff,
                                                    const float* __restrict__ denom, int64_t dstride,
                                                    const int* __restrict__ done) {
  if (done && *done) return;
  const int b = blockIdx.y;
  const float d = denom[b * dstride];
  float* col = V + b * ldb + coloff;
  if (V4) {
    const int64_t i = 4 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (i + 3 < n) {
      float4 v = *reinterpret_cast<const float4*>(col + i);
      v.x /= d;
      v.y /= d;
      v.z /= d;
      v.w /= d;
      *reinterpret_cast<float4*>(col + i) = v;
    } else {
      for (int64_t k = i; k < n; ++k) col[k] /= d;
    }
  } else {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    col[i] /= d;
  }
}

inline void launch_scale_col(int64_t n, int nb, float* V, int64_t ldb, int64_t coloff, const float* denom,
                             int64_t dstride, const int* done, hipStream_t s) {
  if (ldb % 4 == 0 && coloff % 4 == 0 && (reinterpret_cast<uintptr_t>(V) & 15) == 0)
    hipLaunchKernelGGL(lz_scale_col<true>, dim3((unsigned)cdiv(cdiv(n, 4), 256), nb), dim3(256), 0, s, n, V, ldb,
                       coloff, denom, dstride, done);
  else
    hipLaunchKernelGGL(lz_scale_col<false>, dim3((unsigned)cdiv(n, 256), nb), dim3(256), 0, s, n, V, ldb, coloff,
                       denom, dstride, done);
}

// Projection partials h_j = V[:,j] . w over j in [j0, j1] for one tile.
// FIRST (Lanczos): alpha = sum(dpart) (Pd partials per probe), w <- w - (alpha v_i + beta v_{i-1})
// first and alpha is written to diag[b][i-1].
// elements idx..idx+3 of p, zero past n (16-byte load when all four are in range)
__device__ __forceinline__ float4 load4_masked(const float* __restrict__ p, int64_t idx, int64_t n) {
  if (idx + 3 < n) return *reinterpret_cast<const float4*>(p + idx);
  float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
  if (idx < n) r.x = p[idx];
  if (idx + 1 < n) r.y = p[idx + 1];
  if (idx + 2 < n) r.z = p[idx + 2];
  return r;
}

// elements idx..idx+3 of p: a plain 16-byte load when the caller knows they are all below n
// (every tile but a vector's last), so no load sits behind a bounds branch that would make
// the compiler wait for the loads before it.  (The two-sweep kernels keep the masked loads:
// there the plain form measured 11-17% slower on the 1024^2 Laplacian basis.)
template <bool FULL>
__device__ __forceinline__ float4 load4(const float* __restrict__ p, int64_t idx, int64_t n) {
  if constexpr (FULL) return *reinterpret_cast<const float4*>(p + idx);
  else return load4_masked(p, idx, n);
}

// hpart[j * P + blk] = V[:, j]^T w over this block's 1024-element tile, j0 <= j <= j1, with
// the tile of w staged in wl.  Each lane owns 4 consecutive elements of each 256-element
// quarter; two basis vectors (j, j+4) per pass keep eight 16-byte loads in flight per lane.
// REV walks the columns from j1 down (the fused kernel: the columns its update read last
// are the likeliest still in L2); each column's sum is the same either way.
template <bool REV = false>
__device__ __forceinline__ void project_tile_vec(const float* __restrict__ Vb, int64_t ldv, int64_t n, int64_t i0,
                                                 int j0, int j1, const float* wl, float* __restrict__ hpart, int P,
                                                 int blk) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float4* wl4 = reinterpret_cast<const float4*>(wl);
  float4 wv[kTile / 256];
#pragma unroll
  for (int q = 0; q < kTile / 256; ++q) wv[q] = wl4[q * 64 + lane];
  for (int jj = j0 + wave; jj <= j1; jj += 8) {
    // column pair (j, j + 4), or (j1 + j0 - jj, that - 4) walking down
    const int j = REV ? j1 + j0 - jj : jj;
    const int j2 = REV ? j - 4 : j + 4;
    const bool two = jj + 4 <= j1;
    const float* v0 = Vb + (int64_t)j * ldv;
    const float* v1 = Vb + (int64_t)j2 * ldv;
    float a0 = 0.f, a1 = 0.f;
#pragma unroll
    for (int q = 0; q < kTile / 256; ++q) {
      const int64_t idx = i0 + 256 * q + 4 * lane;
      const float4 x0 = load4_masked(v0, idx, n);
      const float4 x1 = two ? load4_masked(v1, idx, n) : make_float4(0.f, 0.f, 0.f, 0.f);
      a0 = fmaf(x0.x, wv[q].x, a0);
      a0 = fmaf(x0.y, wv[q].y, a0);
      a0 = fmaf(x0.z, wv[q].z, a0);
      a0 = fmaf(x0.w, wv[q].w, a0);
      a1 = fmaf(x1.x, wv[q].x, a1);
      a1 = fmaf(x1.y, wv[q].y, a1);
      a1 = fmaf(x1.z, wv[q].z, a1);
      a1 = fmaf(x1.w, wv[q].w, a1);
    }
    a0 = wave_sum(a0);
    a1 = wave_sum(a1);
    if (lane == 0) {
      hpart[(int64_t)j * P + blk] = a0;
      if (two) hpart[(int64_t)j2 * P + blk] = a1;
    }
  }
}

// sum_{j0<=j<=j1} V[idx .. idx+3, j] h[j] for one thread's 4 consecutive elements; four
// basis vectors per pass (sixteen loads in flight)
__device__ __forceinline__ float4 apply_sum_vec(const float* __restrict__ Vb, int64_t ldv, int64_t n, int64_t idx,
                                                int j0, int j1, const float* hs) {
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  int j = j0;
  for (; j + 3 <= j1; j += 4) {
    float4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = load4_masked(Vb + (int64_t)(j + u) * ldv, idx, n);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float hj = hs[j + u];
      acc.x = fmaf(v[u].x, hj, acc.x);
      acc.y = fmaf(v[u].y, hj, acc.y);
      acc.z = fmaf(v[u].z, hj, acc.z);
      acc.w = fmaf(v[u].w, hj, acc.w);
    }
  }
  for (; j <= j1; ++j) {
    const float4 v = load4_masked(Vb + (int64_t)j * ldv, idx, n);
    const float hj = hs[j];
    acc.x = fmaf(v.x, hj, acc.x);
    acc.y = fmaf(v.y, hj, acc.y);
    acc.z = fmaf(v.z, hj, acc.z);
    acc.w = fmaf(v.w, hj, acc.w);
  }
  return acc;
}

// w[idx .. idx+3] -= acc (elements below n); returns the new values (zero past n)
__device__ __forceinline__ float4 apply_store_vec(float* __restrict__ w, int64_t n, int64_t idx, float4 acc) {
  if (idx + 3 < n) {
    float4 wv = *reinterpret_cast<const float4*>(w + idx);
    wv.x -= acc.x;
    wv.y -= acc.y;
    wv.z -= acc.z;
    wv.w -= acc.w;
    *reinterpret_cast<float4*>(w + idx) = wv;
    return wv;
  }
  const float a4[4] = {acc.x, acc.y, acc.z, acc.w};
  float r[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (idx + q < n) {
      r[q] = w[idx + q] - a4[q];
      w[idx + q] = r[q];
    }
  }
  return make_float4(r[0], r[1], r[2], r[3]);
}

template <bool FIRST, bool VEC>
__global__ __launch_bounds__(256) void lz_project(int64_t n, float* __restrict__ V, int64_t ldv, int64_t ldb, int i,
                                                  int j0, int j1, const float* __restrict__ dpart, int Pd, int P,
                                                  float* __restrict__ hpart, int M2, float* __restrict__ diag,
                                                  const float* __restrict__ subdiag, int m,
                                                  const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ float wl[kTile];
  __shared__ float coef[2];
  const int b = blockIdx.y, blk = blockIdx.x;
  const int64_t i0 = (int64_t)blk * kTile;
  float* Vb = V + b * ldb;
  float* w = Vb + (int64_t)(i + 1) * ldv;
  if (FIRST) {
    if (threadIdx.x < 64) {
      const double a = col_total(dpart, Pd, (int64_t)b * Pd);
      if (threadIdx.x == 0) {
        coef[0] = (float)a;
        coef[1] = subdiag[b * (m + 1) + (i - 1)];
        if (blk == 0) diag[b * m + (i - 1)] = (float)a;
      }
    }
    __syncthreads();
  }
  const float alpha = FIRST ? coef[0] : 0.f;
  const float betap = FIRST ? coef[1] : 0.f;
  for (int e = threadIdx.x; e < kTile; e += 256) {
    const int64_t idx = i0 + e;
    float wv = 0.f;
    if (idx < n) {
      wv = w[idx];
      if (FIRST) {
        // lanczos.py:247-248: aux = alpha v_i + beta v_{i-1}; w -= aux
        const float aux = alpha * Vb[(int64_t)i * ldv + idx] + betap * Vb[(int64_t)(i - 1) * ldv + idx];
        wv = wv - aux;
        w[idx] = wv;
      }
    }
    wl[e] = wv;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (VEC) {
    project_tile_vec(Vb, ldv, n, i0, j0, j1, wl, hpart + (int64_t)b * M2 * P, P, blk);
  } else {
    for (int j = j0 + wave; j <= j1; j += 4) {
      const float* vj = Vb + (int64_t)j * ldv;
      float acc = 0.f;
      for (int e = lane; e < kTile; e += 64) {
        const int64_t idx = i0 + e;
        if (idx < n) acc = fmaf(vj[idx], wl[e], acc);
      }
      acc = wave_sum(acc);
      if (lane == 0) hpart[((int64_t)b * M2 + j) * P + blk] = acc;
    }
  }
}


// h[b][j] = sum_p hpart[b][j][p] for j in [j0, j1]; ACCUM adds into h2 (Arnoldi CGS2 keeps h1+h2)
__global__ __launch_bounds__(256) void lz_reduce_h(const float* __restrict__ hpart, int P, int M2, int j0, int j1,
                                                   float* __restrict__ h, const int* __restrict__ done) {
  if (done && *done) return;
  const int b = blockIdx.y;
  const int j = j0 + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j > j1) return;
  const double s = col_total(hpart, P, ((int64_t)b * M2 + j) * P);
  if ((threadIdx.x & 63) == 0) h[b * M2 + j] = (float)s;
}

// w -= sum_{j0<=j<=j1} V[:,j] h[j]  (the sum is formed first, as lanczos.py:295).
// NORM: also writes partials of ||w||^2.
template <bool NORM, bool VEC>
__global__ __launch_bounds__(256) void lz_apply(int64_t n, float* __restrict__ V, int64_t ldv, int64_t ldb, int wcol,
                                                int j0, int j1, const float* __restrict__ h, int M2,
                                                float* __restrict__ npart, int P, const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ float hs[1056];
  __shared__ float red[4];
  const int b = blockIdx.y, blk = blockIdx.x;
  for (int j = threadIdx.x; j < M2 && j < 1056; j += 256) hs[j] = h[b * M2 + j];
  __syncthreads();
  float* Vb = V + b * ldb;
  float* w = Vb + (int64_t)wcol * ldv;
  const int64_t i0 = (int64_t)blk * kTile;
  float ss = 0.f;
  if (VEC) {
    // 4 consecutive elements per thread
    const int64_t idx = i0 + 4 * threadIdx.x;
    const float4 wv = apply_store_vec(w, n, idx, apply_sum_vec(Vb, ldv, n, idx, j0, j1, hs));
    // elements past n are zero, so the tail adds nothing
    ss = fmaf(wv.x, wv.x, ss);
    ss = fmaf(wv.y, wv.y, ss);
    ss = fmaf(wv.z, wv.z, ss);
    ss = fmaf(wv.w, wv.w, ss);
  } else {
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int j = j0; j <= j1; ++j) {
      const float hj = hs[j];
      const float* vj = Vb + (int64_t)j * ldv;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t idx = i0 + threadIdx.x + 256 * q;
        if (idx < n) acc[q] = fmaf(vj[idx], hj, acc[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t idx = i0 + threadIdx.x + 256 * q;
      if (idx < n) {
        const float wv = w[idx] - acc[q];
        w[idx] = wv;
        ss = fmaf(wv, wv, ss);
      }
    }
  }
  if (NORM) {
    ss = block_sum<256>(ss, red);
    if (threadIdx.x == 0) npart[(int64_t)b * P + blk] = ss;
  }
}

// lz_apply<false, true> then lz_project<false, true> in one launch: w -= V[:, j0..j1] h,
// then hpart = V[:, j0..j1]^T w for the updated w (the second pass of double Gram-Schmidt
// right after the first pass's update).  The block's updated tile of w goes to LDS, so
// the projection's arithmetic is the unfused kernel's; its basis rows are the ones the
// update just read, served from the caches instead of HBM.
__global__ __launch_bounds__(256) void lz_apply_project(int64_t n, float* __restrict__ V, int64_t ldv, int64_t ldb,
                                                        int wcol, int j0, int j1, const float* __restrict__ h,
                                                        int M2, float* __restrict__ hpart, int P,
                                                        const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ float hs[1056];
  __shared__ __align__(16) float wl[kTile];
  const int b = blockIdx.y, blk = blockIdx.x;
  for (int j = threadIdx.x; j < M2 && j < 1056; j += 256) hs[j] = h[b * M2 + j];
  __syncthreads();
  float* Vb = V + b * ldb;
  float* w = Vb + (int64_t)wcol * ldv;
  const int64_t i0 = (int64_t)blk * kTile;
  const int64_t idx = i0 + 4 * threadIdx.x;
  reinterpret_cast<float4*>(wl)[threadIdx.x] = apply_store_vec(w, n, idx, apply_sum_vec(Vb, ldv, n, idx, j0, j1, hs));
  __syncthreads();
  project_tile_vec<true>(Vb, ldv, n, i0, j0, j1, wl, hpart + (int64_t)b * M2 * P, P, blk);
}

// RPL consecutive elements idx.. of p (zero past n): one 16-, 8- or 4-byte load
template <int RPL, bool FULL>
__device__ __forceinline__ void load_rows(const float* __restrict__ p, int64_t idx, int64_t n, float (&r)[RPL]) {
  if constexpr (RPL == 4) {
    const float4 v = load4<FULL>(p, idx, n);
    r[0] = v.x, r[1] = v.y, r[2] = v.z, r[3] = v.w;
  } else if constexpr (RPL == 2) {
    if (FULL || idx + 1 < n) {
      const float2 v = *reinterpret_cast<const float2*>(p + idx);
      r[0] = v.x, r[1] = v.y;
    } else {
      r[0] = idx < n ? p[idx] : 0.f;
      r[1] = 0.f;
    }
  } else {
    r[0] = (FULL || idx < n) ? p[idx] : 0.f;
  }
}

// The same two passes as lz_apply_project with one read of the basis: the block's 1024-row
// tile is walked in sub-tiles of SR rows; each wave takes the columns j0 + wave + 4c
// (c < 4 JW, so up to 16 JW columns; 32 KB of LDS for the staged rows), adds their share of
// V h for its lane's RPL = SR / 64 rows and keeps the loaded rows in LDS; after one barrier
// every wave forms the updated w for its lane's rows (the four shares in wave order), wave 0
// stores it, and each wave accumulates its columns' dot products from the LDS copy.  The
// per-lane sums run across the sub-tiles; one wave sum per column at the end gives the
// tile's partial.  The tiles below n (all but a vector's last) load without bounds masks.
template <int SR, int JW, bool FULL>
__device__ __forceinline__ void apply_project_tile(int64_t n, const float* __restrict__ Vb, int64_t ldv,
                                                   float* __restrict__ w, int64_t i0, int j0, int j1,
                                                   const float* hs, float (*Vs)[SR], float (*red)[SR],
                                                   float* __restrict__ hp, int P, int blk) {
  constexpr int RPL = SR / 64;
  constexpr int NC = 4 * JW;  // column slots per wave
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // columns of this wave: j0 + wv + 4 c for c < nc
  const int nc = j0 + wv > j1 ? 0 : (j1 - j0 - wv) / 4 + 1;
  float acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0.f;
  for (int st = 0; st < kTile / SR; ++st) {
    const int64_t idx = i0 + st * SR + RPL * lane;
    float wr[RPL], part[RPL];
    load_rows<RPL, FULL>(w, idx, n, wr);
#pragma unroll
    for (int q = 0; q < RPL; ++q) part[q] = 0.f;
    // four columns per batch, loads clamped to the last column and issued together
#pragma unroll
    for (int c0 = 0; c0 < NC; c0 += 4) {
      if (c0 >= nc) break;
      float v[4][RPL];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        load_rows<RPL, FULL>(Vb + (int64_t)min(j0 + wv + 4 * (c0 + u), j1) * ldv, idx, n, v[u]);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        float* dst = &Vs[wv * NC + c0 + u][RPL * lane];
#pragma unroll
        for (int q = 0; q < RPL; ++q) dst[q] = v[u][q];
        const float hv = hs[min(j0 + wv + 4 * (c0 + u), j1)];
        const float hj = c0 + u < nc ? hv : 0.f;
#pragma unroll
        for (int q = 0; q < RPL; ++q) part[q] = fmaf(v[u][q], hj, part[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < RPL; ++q) red[wv][RPL * lane + q] = part[q];
    __syncthreads();
    float w1[RPL];
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      const int r = RPL * lane + q;
      w1[q] = wr[q] - (((red[0][r] + red[1][r]) + red[2][r]) + red[3][r]);
    }
    if (wv == 0) {
#pragma unroll
      for (int q = 0; q < RPL; ++q)
        if (FULL || idx + q < n) w[idx + q] = w1[q];
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (c >= nc) break;
      const float* src = &Vs[wv * NC + c][RPL * lane];
      float d = 0.f;
#pragma unroll
      for (int q = 0; q < RPL; ++q) d = fmaf(src[q], w1[q], d);
      acc[c] += d;
    }
    __syncthreads();  // Vs and red are rewritten by the next sub-tile
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (c >= nc) break;
    const float sum = wave_sum(acc[c]);
    if (lane == 0) hp[(int64_t)(j0 + wv + 4 * c) * P + blk] = sum;
  }
}

template <int SR, int JW>
__global__ __launch_bounds__(256) void lz_apply_project_lds(int64_t n, float* __restrict__ V, int64_t ldv,
                                                            int64_t ldb, int wcol, int j0, int j1,
                                                            const float* __restrict__ h, int M2,
                                                            float* __restrict__ hpart, int P,
                                                            const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ float hs[1056];
  __shared__ __align__(16) float Vs[16 * JW][SR];
  __shared__ __align__(16) float red[4][SR];
  const int b = blockIdx.y, blk = blockIdx.x;
  for (int j = threadIdx.x; j < M2 && j < 1056; j += 256) hs[j] = h[b * M2 + j];
  __syncthreads();
  float* Vb = V + b * ldb;
  float* w = Vb + (int64_t)wcol * ldv;
  const int64_t i0 = (int64_t)blk * kTile;
  float* hp = hpart + (int64_t)b * M2 * P;
  if (i0 + kTile <= n)
    apply_project_tile<SR, JW, true>(n, Vb, ldv, w, i0, j0, j1, hs, Vs, red, hp, P, blk);
  else
    apply_project_tile<SR, JW, false>(n, Vb, ldv, w, i0, j0, j1, hs, Vs, red, hp, P, blk);
}

// w -= V[:, j0..j1] h, then the partials of V[:, j0..j1]^T w (the two CGS2 passes between
// the reductions): one read of the basis for up to 32 columns, else lz_apply_project.
inline void launch_apply_project(dim3 grid, hipStream_t s, int64_t n, float* V, int64_t ldv, int64_t ldb, int wcol,
                                 int j0, int j1, const float* h, int M2, float* hpart, int P, const int* done) {
  // measured on a 1024^2 Laplacian basis: the staged kernel wins up to 32 columns (SLQ's 32-probe
  // bases of 30 steps: 14% per step); with more columns its smaller sub-tiles and lower occupancy
  // lose to the two-sweep kernel, whose second sweep partly hits the caches
  if (j1 - j0 + 1 <= 32)
    hipLaunchKernelGGL((lz_apply_project_lds<256, 2>), grid, dim3(256), 0, s, n, V, ldv, ldb, wcol, j0, j1, h, M2,
                       hpart, P, done);
  else
    hipLaunchKernelGGL(lz_apply_project, grid, dim3(256), 0, s, n, V, ldv, ldb, wcol, j0, j1, h, M2, hpart, P, done);
}

// beta_i = ||w||, errors and the loop condition (lanczos.py:251-268)
__global__ __launch_bounds__(1024) void lz_finalize(int nb, int m, const float* __restrict__ npart, int P,
                                                    float* __restrict__ subdiag, void* state, float tol, int i) {
  ck_lz_ctrl* c = static_cast<ck_lz_ctrl*>(state);
  if (c->done) return;
  float* errors = reinterpret_cast<float*>(c + 1);
  // each probe's wave forms beta_i, its ratio to beta_1 and its continue flag; thread 0 then
  // takes the maximum and the any over probes in index order from LDS
  __shared__ float s_rel[4096];
  __shared__ unsigned char s_any[4096];
  const int wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int inext = i + 1;
  for (int b = wave; b < nb; b += nwave) {
    const double s = col_total(npart, P, (int64_t)b * P);
    if ((threadIdx.x & 63) == 0) {
      float* sd = subdiag + b * (m + 1);
      const float beta = (float)sqrt(s);
      sd[i] = beta;
      const float beta1 = i == 1 ? beta : sd[1];
      s_rel[b] = beta / fmaxf(beta1, 1e-30f);
      s_any[b] = (beta > tol * beta1) || (inext <= 1);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float err = -INFINITY;
    bool any = false;
    for (int b = 0; b < nb; ++b) {
      err = fmaxf(err, s_rel[b]);
      any = any || s_any[b];
    }
    errors[inext - 1] = err + (inext <= 1 ? 1.f : 0.f);
    c->i = inext;
    if (!((inext <= m) && any)) c->done = 1;
  }
}

// ---------------------------------------------------------------- Arnoldi (MGS)
// One dependent Gram-Schmidt step j on w = Q[:, idx+1]: h_j = (partials of q_j.w),
// w -= h_j q_j, then partials of the next product (q_{j+1}.w, or ||w||^2 when j == idx).
__global__ __launch_bounds__(256) void ar_mgs_step(int64_t n, float* __restrict__ Q, int64_t ldv, int64_t ldb,
                                                   int idx, int j, const float* __restrict__ pin,
                                                   float* __restrict__ pout, int P, float* __restrict__ H, int m,
                                                   const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ float red[4];
  __shared__ float hj_s;
  const int b = blockIdx.y, blk = blockIdx.x;
  if (threadIdx.x < 64) {
    const double hj = col_total(pin, P, (int64_t)b * P);
    if (threadIdx.x == 0) {
      hj_s = (float)hj;
      if (blk == 0) H[(int64_t)b * (m + 1) * m + (int64_t)j * m + idx] = (float)hj;
    }
  }
  __syncthreads();
  const float hj = hj_s;
  float* Qb = Q + b * ldb;
  const float* qj = Qb + (int64_t)j * ldv;
  float* w = Qb + (int64_t)(idx + 1) * ldv;
  const float* qn = (j < idx) ? Qb + (int64_t)(j + 1) * ldv : nullptr;
  const int64_t i0 = (int64_t)blk * kTile;
  float acc = 0.f;
  for (int e = threadIdx.x; e < kTile; e += 256) {
    const int64_t i = i0 + e;
    if (i < n) {
      const float wv = w[i] - hj * qj[i];  // arnoldi.py:308
      w[i] = wv;
      acc = fmaf(qn ? qn[i] : wv, wv, acc);
    }
  }
  acc = block_sum<256>(acc, red);
  if (threadIdx.x == 0) pout[(int64_t)b * P + blk] = acc;
}

// norm = ||w||; H[idx+1, idx] = norm; w /= max(norm, tol/2); loop condition (arnoldi.py:293-319)
__global__ __launch_bounds__(256) void ar_finalize(int nb, int m, const float* __restrict__ npart, int P,
                                                   float* __restrict__ H, float* __restrict__ nrm, void* state,
                                                   float tol, int idx) {
  ck_ar_ctrl* c = static_cast<ck_ar_ctrl*>(state);
  if (c->done) return;
  float* errors = reinterpret_cast<float*>(c + 1);
  const int wave = threadIdx.x >> 6;
  for (int b = wave; b < nb; b += 4) {
    const double s = col_total(npart, P, (int64_t)b * P);
    if ((threadIdx.x & 63) == 0) {
      const float nv = (float)sqrt(s);
      nrm[b] = nv;
      H[(int64_t)b * (m + 1) * m + (int64_t)(idx + 1) * m + idx] = nv;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int inext = idx + 1;
    bool any = false;
    for (int b = 0; b < nb; ++b) {
      const float h10 = H[(int64_t)b * (m + 1) * m + 1 * m + 0];
      any = any || (nrm[b] > tol * h10) || (inext <= 0);
    }
    errors[inext] = nrm[0];
    c->idx = inext;
    if (!((inext < m) && any)) c->done = 1;
  }
}

// w /= max(norm, tol/2); runs for the step that just finished even if it set done.
__global__ __launch_bounds__(256) void ar_normalize(int64_t n, float* __restrict__ Q, int64_t ldv, int64_t ldb,
                                                    int col, const float* __restrict__ nrm, float clipv,
                                                    const int* __restrict__ ctrl_idx, int expect_idx) {
  if (*ctrl_idx != expect_idx) return;  // the step did not run
  const int b = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Q[b * ldb + (int64_t)col * ldv + i] /= fmaxf(nrm[b], clipv);
}

// Arnoldi CGS2: copy h1 + h2 into H column idx.
__global__ __launch_bounds__(256) void ar_store_h(int nb, int m, int M2, int idx, const float* __restrict__ h1,
                                                  const float* __restrict__ h2, float* __restrict__ H,
                                                  const int* __restrict__ done) {
  if (done && *done) return;
  const int b = blockIdx.x;
  for (int j = threadIdx.x; j <= idx; j += 256)
    H[(int64_t)b * (m + 1) * m + (int64_t)j * m + idx] = h1[b * M2 + j] + h2[b * M2 + j];
}

__global__ __launch_bounds__(256) void ar_start_ctrl(int nb, int m, float* __restrict__ H, const float* __restrict__ nrm,
                                                     void* state) {
  ck_ar_ctrl* c = static_cast<ck_ar_ctrl*>(state);
  float* errors = reinterpret_cast<float*>(c + 1);
  for (int64_t t = threadIdx.x; t < (int64_t)nb * (m + 1) * m; t += 256) H[t] = 0.f;
  if (threadIdx.x == 0) {
    c->idx = 0;
    c->max_iters = m;
    c->nb = nb;
    c->done = (0 < m) ? 0 : 1;
    errors[0] = nrm[0];
  }
}

// 16-byte basis access needs 4-float aligned columns and probe blocks
inline bool basis_vec4(const float* V, int64_t ldv, int64_t ldb) {
  return ldv % 4 == 0 && ldb % 4 == 0 && (reinterpret_cast<uintptr_t>(V) & 15) == 0;
}

template <bool FIRST, typename... Args>
void launch_project(bool vec, dim3 grid, hipStream_t s, Args... args) {
  if (vec)
    hipLaunchKernelGGL((lz_project<FIRST, true>), grid, dim3(256), 0, s, args...);
  else
    hipLaunchKernelGGL((lz_project<FIRST, false>), grid, dim3(256), 0, s, args...);
}

template <bool NORM, typename... Args>
void launch_apply(bool vec, dim3 grid, hipStream_t s, Args... args) {
  if (vec)
    hipLaunchKernelGGL((lz_apply<NORM, true>), grid, dim3(256), 0, s, args...);
  else
    hipLaunchKernelGGL((lz_apply<NORM, false>), grid, dim3(256), 0, s, args...);
}

int check_basis(int64_t n, int nb, const float* V, int64_t ldv, int64_t ldb, int m, int extra) {
  if (n < 1 || nb < 1 || nb > 4096 || !V || ldv < n || m < 1 || ldb < (int64_t)(m + extra) * ldv)
    return CK_ERR_BAD_ARG;
  if (m + 2 > 1056) return CK_ERR_UNSUPPORTED;
  return CK_OK;
}

// Q[b][0][i] = s[b][i] / nrm[b]
__global__ __launch_bounds__(256) void ar_start_fill(int64_t n, const float* __restrict__ s, int64_t lds_,
                                                     float* __restrict__ Q, int64_t ldb,
                                                     const float* __restrict__ nrm) {
  const int b = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Q[b * ldb + i] = s[b * lds_ + i] / nrm[b];
}

int lz_pre_full(int64_t n, int nb, float* V, int64_t ldv, int64_t ldb, const float* subdiag, void* state, int m,
                int step, void* scratch, hipStream_t s) {
  const int i = step + 1;
  const int* done = &static_cast<ck_lz_ctrl*>(state)->done;
  LzScratch sc = scratch_of(scratch, n, nb, m);
  if (step == 0) {
    // the reference re-normalises column 1 at the first step (lanczos.py:240-241)
    hipLaunchKernelGGL(tile_partials<1>, dim3(sc.P, nb), dim3(256), 0, s, n, V + ldv, ldb, V + ldv, ldb, sc.dpart,
                       sc.P, done);
    hipLaunchKernelGGL(lz_norm_final, dim3(1), dim3(256), 0, s, nb, sc.dpart, sc.P, sc.nrm, done);
    launch_scale_col(n, nb, V, ldb, ldv, (const float*)sc.nrm, (int64_t)1, done, s);
  } else {
    // V[:, i] was w_{i-1}; its norm is subdiag[i-1] (lanczos.py:252), the same reduction
    launch_scale_col(n, nb, V, ldb, (int64_t)i * ldv, subdiag + (i - 1), (int64_t)(m + 1), done, s);
  }
  CK_CHECK_LAUNCH();
  return CK_OK;
}

// xy (optional): the v_i . w partials the matvec wrote (pxy per probe) instead of a tile_partials pass
int lz_post_full(int64_t n, int nb, float* V, int64_t ldv, int64_t ldb, float* diag, float* subdiag, void* state,
                 int m, float tol, int step, void* scratch, hipStream_t s, const float* xy = nullptr, int pxy = 0) {
  const int i = step + 1;
  if (i > m) return CK_ERR_BAD_ARG;
  const int* done = &static_cast<ck_lz_ctrl*>(state)->done;
  LzScratch sc = scratch_of(scratch, n, nb, m);
  const float* w = V + (int64_t)(i + 1) * ldv;
  const float* vi = V + (int64_t)i * ldv;
  dim3 grid(sc.P, nb);
  const bool vec4 = basis_vec4(V, ldv, ldb);
  const float* dp = xy ? xy : sc.dpart;
  const int pd = xy ? pxy : sc.P;
  if (!xy) hipLaunchKernelGGL(tile_partials<0>, grid, dim3(256), 0, s, n, w, ldb, vi, ldb, sc.dpart, sc.P, done);
  launch_project<true>(vec4, grid, s, n, V, ldv, ldb, i, 1, i, dp, pd, sc.P, sc.hpart, sc.M2, diag, subdiag, m,
                       done);
  const dim3 rgrid((unsigned)cdiv(i, 4), nb);
  hipLaunchKernelGGL(lz_reduce_h, rgrid, dim3(256), 0, s, sc.hpart, sc.P, sc.M2, 1, i, sc.h, done);
  if (vec4) {
    launch_apply_project(grid, s, n, V, ldv, ldb, i + 1, 1, i, (const float*)sc.h, sc.M2, sc.hpart, sc.P, done);
  } else {
    launch_apply<false>(vec4, grid, s, n, V, ldv, ldb, i + 1, 1, i, sc.h, sc.M2, sc.dpart, sc.P, done);
    launch_project<false>(vec4, grid, s, n, V, ldv, ldb, i, 1, i, sc.dpart, sc.P, sc.P, sc.hpart, sc.M2, diag,
                          subdiag, m, done);
  }
  hipLaunchKernelGGL(lz_reduce_h, rgrid, dim3(256), 0, s, sc.hpart, sc.P, sc.M2, 1, i, sc.h, done);
  launch_apply<true>(vec4, grid, s, n, V, ldv, ldb, i + 1, 1, i, sc.h, sc.M2, sc.dpart,
                     sc.P, done);
  hipLaunchKernelGGL(lz_finalize, dim3(1), dim3(1024), 0, s, nb, m, sc.dpart, sc.P, subdiag, state, tol, i);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

// ---------------------------------------------------------------- block re-orthogonalisation
// Opt-in mode of ck_lanczos_iterate_block (no counterpart in the reference, whose Lanczos
// re-orthogonalises every new vector against the whole basis, lanczos.py:287-296).  The
// basis is cut into blocks of S columns.  Each step orthogonalises w (CGS2) only against the
// columns of its own block and the column before it; when a block is complete it is made
// orthogonal to all older columns at once, two classical passes of
//   H = V_old^T W  (blk_project)   and   W -= V_old H  (blk_apply)
// on v_mfma_f32_16x16x4f32.  V_old streams from HBM once per pass for S columns instead of
// three times per column, so the basis traffic drops by about 3S/4; between two block
// passes the new columns carry only the components a few recurrence steps let grow.
constexpr int kBlkJ = 256;        // old columns per blk_project / blk_apply launch
constexpr int kBlkTile = 4096;    // vector elements per blk_project block
typedef float f32x4v __attribute__((ext_vector_type(4)));

// hpart[b][(jj * 16 + c) * P2 + tile] = sum over the tile of V[j0 + jj] . V[w0 + c] for
// jj < nj, c < S.  A lane of a 16x16x4 step holds operand k = 16 (lane / 16) + ks of the
// 64-element chunk (the sum over k is order-free), so its A and B values are 16 contiguous
// floats of one column: four float4 loads each.
__global__ __launch_bounds__(256) void blk_project(int64_t ld, const float* __restrict__ V, int64_t ldv,
                                                   int64_t ldb, int j0, int nj, int w0, int S,
                                                   float* __restrict__ hpart, int P2, const int* __restrict__ done) {
  if (done && *done) return;
  const int b = blockIdx.y, tile = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float* Vb = V + b * ldb;
  const int64_t i_begin = (int64_t)tile * kBlkTile;
  const int64_t i_end = i_begin + kBlkTile < ld ? i_begin + kBlkTile : ld;
  const int njt = (nj + 15) / 16;
  const int cl = lane & 15, slot = lane >> 4;
  const float* wcol = Vb + (int64_t)(w0 + (cl < S ? cl : S - 1)) * ldv + 16 * slot;
  const float* vcol[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int jt = wave + 4 * t;
    const int jj = 16 * jt + cl;
    vcol[t] = Vb + (int64_t)(j0 + (jj < nj ? jj : nj - 1)) * ldv + 16 * slot;
  }
  f32x4v acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = f32x4v{0.f, 0.f, 0.f, 0.f};
  for (int64_t i0 = i_begin; i0 < i_end; i0 += 64) {
    float4 wv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) wv[q] = *reinterpret_cast<const float4*>(wcol + i0 + 4 * q);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (wave + 4 * t >= njt) break;
      float4 av[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) av[q] = *reinterpret_cast<const float4*>(vcol[t] + i0 + 4 * q);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q].x, wv[q].x, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q].y, wv[q].y, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q].z, wv[q].z, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q].w, wv[q].w, acc[t], 0, 0, 0);
      }
    }
  }
  // D[m][n]: lane holds m = 4 (lane / 16) + v (old column within the tile), n = lane % 16
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int jt = wave + 4 * t;
    if (jt >= njt) break;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int jj = 16 * jt + 4 * slot + v;
      if (jj < nj && cl < S) hpart[((int64_t)b * kBlkJ * 16 + jj * 16 + cl) * P2 + tile] = acc[t][v];
    }
  }
}

// H[b][jj * 16 + c] = sum over tiles (fixed order, FP64) for jj < nj, c < S; zero for c >= S
__global__ __launch_bounds__(256) void blk_reduce(const float* __restrict__ hpart, int P2, int nj, int S,
                                                  float* __restrict__ H, const int* __restrict__ done) {
  if (done && *done) return;
  const int b = blockIdx.y;
  const int e = blockIdx.x * 4 + (threadIdx.x >> 6);  // (jj, c) entry
  if (e >= nj * 16) return;
  const int c = e % 16;
  double s = 0.0;
  if (c < S) s = col_total(hpart, P2, ((int64_t)b * kBlkJ * 16 + e) * P2);
  if ((threadIdx.x & 63) == 0) H[(int64_t)b * kBlkJ * 16 + e] = (float)s;
}

// W[c] -= sum_{jj < nj} V[j0 + jj] H[jj][c] over 64 elements per wave: a lane of a 16x16x4
// step holds V[j0 + 4 ks + lane / 16][i0 + 4 (lane % 16) + q] as A (row m = lane % 16 of
// sub-tile q, i.e. element 4 m + q) and H[4 ks + lane / 16][lane % 16] as B; one float4 load
// per lane per step serves the four sub-tiles.
__global__ __launch_bounds__(256) void blk_apply(int64_t ld, float* __restrict__ V, int64_t ldv, int64_t ldb, int j0,
                                                 int nj, int w0, int S, const float* __restrict__ H,
                                                 const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ float Hs[kBlkJ][16];
  const int b = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int e = threadIdx.x; e < nj * 16; e += 256) Hs[e / 16][e % 16] = H[(int64_t)b * kBlkJ * 16 + e];
  __syncthreads();
  float* Vb = V + b * ldb;
  const int64_t i0 = (int64_t)blockIdx.x * 256 + 64 * wave;
  if (i0 >= ld) return;
  const int m = lane & 15, slot = lane >> 4;
  f32x4v acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = f32x4v{0.f, 0.f, 0.f, 0.f};
  const int nks = (nj + 3) / 4;
  int ks = 0;
  for (; ks + 4 <= nks; ks += 4) {
    float4 a[4];
    float hb[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int jj = 4 * (ks + u) + slot;
      a[u] = *reinterpret_cast<const float4*>(Vb + (int64_t)(j0 + (jj < nj ? jj : nj - 1)) * ldv + i0 + 4 * m);
      hb[u] = jj < nj ? Hs[jj][m] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].x, hb[u], acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].y, hb[u], acc[1], 0, 0, 0);
      acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].z, hb[u], acc[2], 0, 0, 0);
      acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].w, hb[u], acc[3], 0, 0, 0);
    }
  }
  for (; ks < nks; ++ks) {
    const int jj = 4 * ks + slot;
    const float4 a = *reinterpret_cast<const float4*>(Vb + (int64_t)(j0 + (jj < nj ? jj : nj - 1)) * ldv + i0 + 4 * m);
    const float hb = jj < nj ? Hs[jj][m] : 0.f;
    acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, hb, acc[0], 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, hb, acc[1], 0, 0, 0);
    acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.z, hb, acc[2], 0, 0, 0);
    acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.w, hb, acc[3], 0, 0, 0);
  }
  // sub-tile q, D[m'][n] at lane: m' = 4 slot + v (element i0 + 4 m' + q), n = lane % 16 (column)
  const int c = m;
  if (c >= S) return;
  float* wc = Vb + (int64_t)(w0 + c) * ldv + i0;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int base = 4 * (4 * slot + v);
    float4 wv = *reinterpret_cast<float4*>(wc + base);
    wv.x -= acc[0][v];
    wv.y -= acc[1][v];
    wv.z -= acc[2][v];
    wv.w -= acc[3][v];
    *reinterpret_cast<float4*>(wc + base) = wv;
  }
}

struct BlkScratch {
  float* hpart;  // nb x kBlkJ x 16 x P2
  float* H;      // nb x kBlkJ x 16
  int P2;
};

inline BlkScratch blk_scratch_of(void* scratch, int64_t ld, int nb) {
  BlkScratch s;
  s.P2 = (int)cdiv(ld, kBlkTile);
  s.hpart = static_cast<float*>(scratch);
  s.H = s.hpart + (size_t)nb * kBlkJ * 16 * s.P2;
  return s;
}

// columns [w0, w0 + S) orthogonal to columns [1, w0): two classical passes, old columns in
// chunks of kBlkJ
int blk_orthogonalize(int64_t ld, int nb, float* V, int64_t ldv, int64_t ldb, int w0, int S, const BlkScratch& bs,
                      const int* done, hipStream_t s) {
  for (int pass = 0; pass < 2; ++pass) {
    for (int j0 = 1; j0 < w0; j0 += kBlkJ) {
      const int nj = w0 - j0 < kBlkJ ? w0 - j0 : kBlkJ;
      hipLaunchKernelGGL(blk_project, dim3(bs.P2, nb), dim3(256), 0, s, ld, (const float*)V, ldv, ldb, j0, nj, w0, S,
                         bs.hpart, bs.P2, done);
      hipLaunchKernelGGL(blk_reduce, dim3((unsigned)cdiv(nj * 16, 4), nb), dim3(256), 0, s, (const float*)bs.hpart,
                         bs.P2, nj, S, bs.H, done);
      hipLaunchKernelGGL(blk_apply, dim3((unsigned)cdiv(ld, 256), nb), dim3(256), 0, s, ld, V, ldv, ldb, j0, nj, w0,
                         S, (const float*)bs.H, done);
    }
  }
  CK_CHECK_LAUNCH();
  return CK_OK;
}

// lz_post_full with the re-orthogonalisation restricted to columns [jlo, i]
int lz_post_window(int64_t n, int nb, float* V, int64_t ldv, int64_t ldb, float* diag, float* subdiag, void* state,
                   int m, float tol, int step, int jlo, void* scratch, hipStream_t s, const float* xy = nullptr,
                   int pxy = 0) {
  const int i = step + 1;
  if (i > m) return CK_ERR_BAD_ARG;
  const int* done = &static_cast<ck_lz_ctrl*>(state)->done;
  LzScratch sc = scratch_of(scratch, n, nb, m);
  const float* w = V + (int64_t)(i + 1) * ldv;
  const float* vi = V + (int64_t)i * ldv;
  dim3 grid(sc.P, nb);
  const bool vec4 = basis_vec4(V, ldv, ldb);
  const float* dp = xy ? xy : sc.dpart;
  const int pd = xy ? pxy : sc.P;
  if (!xy) hipLaunchKernelGGL(tile_partials<0>, grid, dim3(256), 0, s, n, w, ldb, vi, ldb, sc.dpart, sc.P, done);
  launch_project<true>(vec4, grid, s, n, V, ldv, ldb, i, jlo, i, dp, pd, sc.P, sc.hpart, sc.M2, diag, subdiag, m,
                       done);
  const dim3 rgrid((unsigned)cdiv(i - jlo + 1, 4), nb);
  hipLaunchKernelGGL(lz_reduce_h, rgrid, dim3(256), 0, s, sc.hpart, sc.P, sc.M2, jlo, i, sc.h, done);
  if (vec4) {
    launch_apply_project(grid, s, n, V, ldv, ldb, i + 1, jlo, i, (const float*)sc.h, sc.M2, sc.hpart, sc.P, done);
  } else {
    launch_apply<false>(vec4, grid, s, n, V, ldv, ldb, i + 1, jlo, i, sc.h, sc.M2, sc.dpart, sc.P, done);
    launch_project<false>(vec4, grid, s, n, V, ldv, ldb, i, jlo, i, sc.dpart, sc.P, sc.P, sc.hpart, sc.M2, diag,
                          subdiag, m, done);
  }
  hipLaunchKernelGGL(lz_reduce_h, rgrid, dim3(256), 0, s, sc.hpart, sc.P, sc.M2, jlo, i, sc.h, done);
  launch_apply<true>(vec4, grid, s, n, V, ldv, ldb, i + 1, jlo, i, sc.h, sc.M2, sc.dpart, sc.P, done);
  hipLaunchKernelGGL(lz_finalize, dim3(1), dim3(1024), 0, s, nb, m, sc.dpart, sc.P, subdiag, state, tol, i);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ar_post_full(int64_t n, int nb, float* Q, int64_t ldv, int64_t ldb, float* H, void* state, int m, float tol,
                 int step, int mgs, void* scratch, hipStream_t s) {
  const int idx = step;
  if (idx >= m) return CK_ERR_BAD_ARG;
  const int* done = &static_cast<ck_ar_ctrl*>(state)->done;
  LzScratch sc = scratch_of(scratch, n, nb, m);
  dim3 grid(sc.P, nb);
  const bool vec4 = basis_vec4(Q, ldv, ldb);
  float* npart;
  if (mgs) {
    float* buf[2] = {sc.dpart, sc.dpart + (size_t)nb * sc.P};
    const float* w = Q + (int64_t)(idx + 1) * ldv;
    hipLaunchKernelGGL(tile_partials<0>, grid, dim3(256), 0, s, n, (const float*)Q, ldb, w, ldb, buf[0], sc.P, done);
    for (int j = 0; j <= idx; ++j)
      hipLaunchKernelGGL(ar_mgs_step, grid, dim3(256), 0, s, n, Q, ldv, ldb, idx, j, (const float*)buf[j & 1],
                         buf[(j + 1) & 1], sc.P, H, m, done);
    npart = buf[(idx + 1) & 1];
  } else {
    const dim3 rgrid((unsigned)cdiv(idx + 1, 4), nb);
    launch_project<false>(vec4, grid, s, n, Q, ldv, ldb, idx, 0, idx, sc.dpart, sc.P, sc.P,
                       sc.hpart, sc.M2, (float*)nullptr, (const float*)nullptr, m, done);
    hipLaunchKernelGGL(lz_reduce_h, rgrid, dim3(256), 0, s, sc.hpart, sc.P, sc.M2, 0, idx, sc.h, done);
    if (vec4) {
      launch_apply_project(grid, s, n, Q, ldv, ldb, idx + 1, 0, idx, (const float*)sc.h, sc.M2, sc.hpart, sc.P,
                           done);
    } else {
      launch_apply<false>(vec4, grid, s, n, Q, ldv, ldb, idx + 1, 0, idx, sc.h, sc.M2, sc.dpart, sc.P, done);
      launch_project<false>(vec4, grid, s, n, Q, ldv, ldb, idx, 0, idx, sc.dpart, sc.P, sc.P, sc.hpart, sc.M2,
                            (float*)nullptr, (const float*)nullptr, m, done);
    }
    hipLaunchKernelGGL(lz_reduce_h, rgrid, dim3(256), 0, s, sc.hpart, sc.P, sc.M2, 0, idx, sc.h2, done);
    launch_apply<true>(vec4, grid, s, n, Q, ldv, ldb, idx + 1, 0, idx, sc.h2, sc.M2,
                       sc.dpart, sc.P, done);
    hipLaunchKernelGGL(ar_store_h, dim3(nb), dim3(256), 0, s, nb, m, sc.M2, idx, (const float*)sc.h,
                       (const float*)sc.h2, H, done);
    npart = sc.dpart;
  }
  hipLaunchKernelGGL(ar_finalize, dim3(1), dim3(256), 0, s, nb, m, (const float*)npart, sc.P, H, sc.nrm, state, tol,
                     idx);
  hipLaunchKernelGGL(ar_normalize, dim3((unsigned)cdiv(n, 256), nb), dim3(256), 0, s, n, Q, ldv, ldb, idx + 1,
                     (const float*)sc.nrm, tol / 2.f, &static_cast<ck_ar_ctrl*>(state)->idx, idx + 1);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // namespace

extern "C" {

size_t ck_lanczos_state_bytes(int32_t nb, int32_t max_iters) {
  return sizeof(ck_lz_ctrl) + sizeof(float) * ((size_t)max_iters + 2);
}

size_t ck_lanczos_scratch_bytes(int64_t n, int32_t nb, int32_t max_iters) {
  const size_t P = lz_nparts(n), M2 = max_iters + 2;
  return sizeof(float) * ((size_t)nb * P * 2 + (size_t)nb * M2 * P + 2 * (size_t)nb * M2 + (size_t)nb + 64);
}

int ck_lanczos_start(int64_t n, int32_t nb, const float* s, int64_t lds_, float* V, int64_t ldv, int64_t ldb,
                     float* diag, float* subdiag, void* state, int32_t max_iters, void* scratch, void* stream) {
  if (int st = check_basis(n, nb, V, ldv, ldb, max_iters, 2)) return st;
  if (!s || lds_ < n || !diag || !subdiag || !state || !scratch) return CK_ERR_BAD_ARG;
  hipStream_t st_ = (hipStream_t)stream;
  LzScratch sc = scratch_of(scratch, n, nb, max_iters);
  hipLaunchKernelGGL(tile_partials<1>, dim3(sc.P, nb), dim3(256), 0, st_, n, s, lds_, s, lds_, sc.dpart, sc.P,
                     (const int*)nullptr);
  hipLaunchKernelGGL(lz_norm_final, dim3(1), dim3(256), 0, st_, nb, sc.dpart, sc.P, sc.nrm, (const int*)nullptr);
  hipLaunchKernelGGL(lz_start_fill, dim3((unsigned)cdiv(n, 256), nb), dim3(256), 0, st_, n, s, lds_, V, ldv, ldb,
                     (const float*)sc.nrm);
  hipLaunchKernelGGL(lz_start_ctrl, dim3(1), dim3(256), 0, st_, nb, max_iters, diag, subdiag, state);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_lanczos_pre(int64_t n, int32_t nb, int32_t max_iters, float* V, int64_t ldv, int64_t ldb,
                   const float* subdiag, void* state, int32_t step, void* scratch, void* stream) {
  if (int st = check_basis(n, nb, V, ldv, ldb, max_iters, 2)) return st;
  if (!subdiag || !state || !scratch || step < 0 || step >= max_iters) return CK_ERR_BAD_ARG;
  return lz_pre_full(n, nb, V, ldv, ldb, subdiag, state, max_iters, step, scratch, (hipStream_t)stream);
}

int ck_lanczos_post(int64_t n, int32_t nb, int32_t max_iters, float* V, int64_t ldv, int64_t ldb, float* diag,
                    float* subdiag, void* state, float tol, int32_t step, void* scratch, void* stream) {
  if (int st = check_basis(n, nb, V, ldv, ldb, max_iters, 2)) return st;
  if (!diag || !subdiag || !state || !scratch || step < 0 || step >= max_iters) return CK_ERR_BAD_ARG;
  return lz_post_full(n, nb, V, ldv, ldb, diag, subdiag, state, max_iters, tol, step, scratch, (hipStream_t)stream);
}

}  // extern "C"

namespace {

// w = A v_i for the nb probes (vector stride ldb); where the operator's kernel has an x . y
// epilogue (CSR with short rows, 3-factor Kronecker) it also returns the v_i . w partials it
// wrote (*xy, *pxy per probe) so the step needs no tile_partials pass
int lz_matvec(const ck_op* A, int nb, const float* vi, float* w, int64_t ldb, void* work, size_t work_bytes,
              const int* done, const LzScratch& sc, const float** xy, int* pxy, hipStream_t s) {
  *xy = nullptr;
  *pxy = 0;
  const int64_t n = A->n;
  if (A->kind == CK_OP_CSR_F32 && work && work_bytes >= csr_xy_bytes(n, nb) && (double)A->nnz <= 32.0 * (double)n) {
    if (int st = csr_apply_xy(A, nb, vi, ldb, w, ldb, done, static_cast<float*>(work), s)) return st;
    *xy = static_cast<const float*>(work);
    *pxy = csr_xy_nparts(n);
    return CK_OK;
  }
  if (A->kind == CK_OP_KRON_F32 && A->nfactors == 3) {
    const int np = kron_nparts(A);
    if (np > 0 && np <= 2 * sc.P) {  // the partials go to dpart (room for nb x 2P)
      const int st = kron_dispatch(A, nb, vi, ldb, w, ldb, work, work_bytes, done, sc.dpart, np, s);
      if (st == CK_OK) {
        *xy = sc.dpart;
        *pxy = np;
        return CK_OK;
      }
      if (st != CK_ERR_UNSUPPORTED) return st;
    }
  }
  return op_apply_internal(A, nb, vi, ldb, w, ldb, work, work_bytes, done, s);
}

}  // namespace

extern "C" {

int ck_lanczos_iterate(const ck_op* A, int64_t n, int32_t nb, int32_t max_iters, float* V, int64_t ldv, int64_t ldb,
                       float* diag, float* subdiag, void* state, float tol, int32_t first_step, int32_t nsteps,
                       void* scratch, void* work, size_t work_bytes, void* stream) {
  if (!A || A->n != n || first_step < 0 || nsteps < 0) return CK_ERR_BAD_ARG;
  if (int st = check_basis(n, nb, V, ldv, ldb, max_iters, 2)) return st;
  if (!diag || !subdiag || !state || !scratch) return CK_ERR_BAD_ARG;
  hipStream_t s = (hipStream_t)stream;
  const int* done = &static_cast<ck_lz_ctrl*>(state)->done;
  const LzScratch sc = scratch_of(scratch, n, nb, max_iters);
  for (int step = first_step; step < first_step + nsteps && step < max_iters; ++step) {
    const int i = step + 1;
    int st = lz_pre_full(n, nb, V, ldv, ldb, subdiag, state, max_iters, step, scratch, s);
    if (st) return st;
    // w = A v_i into column i+1, batched over the nb probes (vector stride ldb)
    const float* xy;
    int pxy;
    st = lz_matvec(A, nb, V + (int64_t)i * ldv, V + (int64_t)(i + 1) * ldv, ldb, work, work_bytes, done, sc, &xy,
                   &pxy, s);
    if (st) return st;
    st = lz_post_full(n, nb, V, ldv, ldb, diag, subdiag, state, max_iters, tol, step, scratch, s, xy, pxy);
    if (st) return st;
  }
  return CK_OK;
}

size_t ck_lanczos_block_scratch_bytes(int64_t n, int32_t nb, int32_t block) {
  if (block < 2 || block > 16 || nb < 1) return 0;
  const int64_t ld = cdiv(n, 64) * 64;
  const size_t P2 = cdiv(ld, kBlkTile);
  return sizeof(float) * ((size_t)nb * kBlkJ * 16 * P2 + (size_t)nb * kBlkJ * 16) + 256;
}

int ck_lanczos_iterate_block(const ck_op* A, int64_t n, int32_t nb, int32_t max_iters, float* V, int64_t ldv,
                             int64_t ldb, float* diag, float* subdiag, void* state, float tol, int32_t first_step,
                             int32_t nsteps, int32_t block, void* scratch, void* block_scratch,
                             size_t block_scratch_bytes, void* work, size_t work_bytes, void* stream) {
  if (!A || A->n != n || first_step < 0 || nsteps < 0 || block < 2 || block > 16) return CK_ERR_BAD_ARG;
  if (int st = check_basis(n, nb, V, ldv, ldb, max_iters, 2)) return st;
  if (!diag || !subdiag || !state || !scratch || !block_scratch) return CK_ERR_BAD_ARG;
  const int64_t ld = cdiv(n, 64) * 64;
  // the block kernels read whole 64-element chunks of every column: zero padding up to ld
  if (ldv < ld || (ldv % 4) || (reinterpret_cast<uintptr_t>(V) & 15)) return CK_ERR_UNSUPPORTED;
  if (block_scratch_bytes < ck_lanczos_block_scratch_bytes(n, nb, block)) return CK_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  const int* done = &static_cast<ck_lz_ctrl*>(state)->done;
  const BlkScratch bs = blk_scratch_of(block_scratch, ld, nb);
  const LzScratch sc = scratch_of(scratch, n, nb, max_iters);
  const int S = block;
  for (int step = first_step; step < first_step + nsteps && step < max_iters; ++step) {
    const int i = step + 1;
    int st = lz_pre_full(n, nb, V, ldv, ldb, subdiag, state, max_iters, step, scratch, s);
    if (st) return st;
    // column i completes block [i - S + 1, i]: orthogonalise it against every older column
    if (i % S == 0 && i > S) {
      st = blk_orthogonalize(ld, nb, V, ldv, ldb, i - S + 1, S, bs, done, s);
      if (st) return st;
    }
    const float* xy;
    int pxy;
    st = lz_matvec(A, nb, V + (int64_t)i * ldv, V + (int64_t)(i + 1) * ldv, ldb, work, work_bytes, done, sc, &xy,
                   &pxy, s);
    if (st) return st;
    // w (column i + 1) against its own block and the column before it
    const int c0 = (i / S) * S + 1;
    const int jlo = c0 - 1 > 1 ? c0 - 1 : 1;
    st = lz_post_window(n, nb, V, ldv, ldb, diag, subdiag, state, max_iters, tol, step, jlo, scratch, s, xy, pxy);
    if (st) return st;
  }
  return CK_OK;
}

size_t ck_arnoldi_state_bytes(int32_t nb, int32_t max_iters) {
  return sizeof(ck_ar_ctrl) + sizeof(float) * ((size_t)max_iters + 2);
}

size_t ck_arnoldi_scratch_bytes(int64_t n, int32_t nb, int32_t max_iters) {
  return ck_lanczos_scratch_bytes(n, nb, max_iters);
}

int ck_arnoldi_start(int64_t n, int32_t nb, const float* s, int64_t lds_, float* Q, int64_t ldv, int64_t ldb, float* H,
                     void* state, int32_t max_iters, void* scratch, void* stream) {
  if (int st = check_basis(n, nb, Q, ldv, ldb, max_iters, 1)) return st;
  if (!s || lds_ < n || !H || !state || !scratch) return CK_ERR_BAD_ARG;
  hipStream_t st_ = (hipStream_t)stream;
  LzScratch sc = scratch_of(scratch, n, nb, max_iters);
  hipLaunchKernelGGL(tile_partials<1>, dim3(sc.P, nb), dim3(256), 0, st_, n, s, lds_, s, lds_, sc.dpart, sc.P,
                     (const int*)nullptr);
  hipLaunchKernelGGL(lz_norm_final, dim3(1), dim3(256), 0, st_, nb, sc.dpart, sc.P, sc.nrm, (const int*)nullptr);
  hipLaunchKernelGGL(ar_start_fill, dim3((unsigned)cdiv(n, 256), nb), dim3(256), 0, st_, n, s, lds_, Q, ldb,
                     (const float*)sc.nrm);
  hipLaunchKernelGGL(ar_start_ctrl, dim3(1), dim3(256), 0, st_, nb, max_iters, H, (const float*)sc.nrm, state);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_arnoldi_post(int64_t n, int32_t nb, int32_t max_iters, float* Q, int64_t ldv, int64_t ldb, float* H,
                    void* state, float tol, int32_t step, int32_t mgs, void* scratch, void* stream) {
  if (int st = check_basis(n, nb, Q, ldv, ldb, max_iters, 1)) return st;
  if (!H || !state || !scratch || step < 0 || step >= max_iters) return CK_ERR_BAD_ARG;
  return ar_post_full(n, nb, Q, ldv, ldb, H, state, max_iters, tol, step, mgs, scratch, (hipStream_t)stream);
}

int ck_arnoldi_iterate(const ck_op* A, int64_t n, int32_t nb, int32_t max_iters, float* Q, int64_t ldv, int64_t ldb,
                       float* H, void* state, float tol, int32_t first_step, int32_t nsteps, int32_t mgs,
                       void* scratch, void* work, size_t work_bytes, void* stream) {
  if (!A || A->n != n || first_step < 0 || nsteps < 0) return CK_ERR_BAD_ARG;
  if (int st = check_basis(n, nb, Q, ldv, ldb, max_iters, 1)) return st;
  if (!H || !state || !scratch) return CK_ERR_BAD_ARG;
  hipStream_t s = (hipStream_t)stream;
  const int* done = &static_cast<ck_ar_ctrl*>(state)->done;
  for (int step = first_step; step < first_step + nsteps && step < max_iters; ++step) {
    int st = op_apply_internal(A, nb, Q + (int64_t)step * ldv, ldb, Q + (int64_t)(step + 1) * ldv, ldb, work,
                               work_bytes, done, s);
    if (st) return st;
    st = ar_post_full(n, nb, Q, ldv, ldb, H, state, max_iters, tol, step, mgs, scratch, s);
    if (st) return st;
  }
  return CK_OK;
}

}  // extern "C"
