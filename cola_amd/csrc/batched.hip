// Batched CG over many independent dense systems of uneven size, one rhs each
// (block-diagonal operators), with per-system storage precision and FP64
// refinement residuals, and HIP-graph capture of the iteration.
//
// Layout.  System s has order n[s]; its matrix is row-major with leading
// dimension lda[s] (a multiple of 8, padding columns zero) at element offset
// moff[s] of the fp32 buffer A32 and, with the same offsets, of the bf16 buffer
// A16.  Its vectors are segments [voff[s], voff[s] + n[s]) of concatenated
// vectors.  Work is listed as tiles: matvec tiles (system, first row) of
// 4 * ROWS rows and vector tiles (system, first index) of kBVec entries, generated
// system by system, so each system's partial sums occupy one contiguous range
// [mt_begin[s], mt_end[s]) / [vt_begin[s], vt_end[s]) and are added in index order.
//
// One CG iteration is five launches for all systems at once:
//   bcg_matvec   Ap = (K_s + shift_s I) p per tile and the tile's p.Ap  [HBM: n^2 bytes x 2 or 4]
//   bcg_alpha    alpha_s = gamma_s / p.Ap                               (one wave per system)
//   bcg_xr       x += alpha p, r -= alpha Ap, tile r.r                   [HBM]
//   bcg_beta     beta_s, gamma_s, k_s, done_s                            (one wave per system)
//   bcg_p        p = r + beta p  (p = z + beta p when preconditioned)    [HBM]
// With a per-system low-rank preconditioner (shift_s I + F_s F_s^T)^-1 three more
// launches per iteration compute z (bcg_pc_ftr, bcg_pc_core, bcg_pc_apply).
// Every kernel returns at once for systems whose done flag is set, so converged
// systems stop costing bandwidth.  The launches contain no host synchronisation and
// no memset, so a run of iterations can be captured into one HIP graph.
//
// The matvec is a GEMV (one rhs): 1 flop per byte of bf16 matrix, far below any MFMA
// ridge, so it runs on the vector ALU; x_s is staged in LDS (4096 entries at a time)
// and shared by the four waves of a tile.
#include "common.hpp"

using namespace ck;

namespace {

constexpr int kBVec = 1024;   // vector tile length
constexpr int kBChunk = 4096; // x entries staged per LDS pass

struct BcgArgs {
  int32_t nsys;
  int32_t max_iters;
  const int32_t* tile_sys;
  const int32_t* tile_row0;
  const int32_t* vtile_sys;
  const int32_t* vtile_i0;
  const int32_t* n;
  const int32_t* lda;
  const int64_t* voff;
  const int64_t* moff;
  const int32_t* mt_begin;
  const int32_t* mt_end;
  const int32_t* vt_begin;
  const int32_t* vt_end;
  const float* shift;
  const int32_t* prec;
  const float* A32;
  const uint16_t* A16;
  float* x;
  float* r;
  float* p;
  float* Ap;
  float* part_mv;
  float* part_vec;
  float* gamma;
  float* alpha;
  float* beta;
  float* thresh;
  int32_t* k;
  int32_t* done;
  int32_t rank;
  const float* F;
  const int64_t* foff;
  const double* Cinv;
  float* z;
  float* part_rz;
  double* part_ftr;
  double* u;
};

BcgArgs to_args(const ck_bcg_plan* P) {
  BcgArgs a;
  a.nsys = P->nsys;
  a.max_iters = P->max_iters;
  a.tile_sys = P->tile_sys;
  a.tile_row0 = P->tile_row0;
  a.vtile_sys = P->vtile_sys;
  a.vtile_i0 = P->vtile_i0;
  a.n = P->n;
  a.lda = P->lda;
  a.voff = P->voff;
  a.moff = P->moff;
  a.mt_begin = P->mt_begin;
  a.mt_end = P->mt_end;
  a.vt_begin = P->vt_begin;
  a.vt_end = P->vt_end;
  a.shift = P->shift;
  a.prec = P->prec;
  a.A32 = P->A32;
  a.A16 = P->A16;
  a.x = P->x;
  a.r = P->r;
  a.p = P->p;
  a.Ap = P->Ap;
  a.part_mv = P->part_mv;
  a.part_vec = P->part_vec;
  a.gamma = P->gamma;
  a.alpha = P->alpha;
  a.beta = P->beta;
  a.thresh = P->thresh;
  a.k = P->k;
  a.done = P->done;
  a.rank = P->rank;
  a.F = P->F;
  a.foff = P->foff;
  a.Cinv = P->Cinv;
  a.z = P->z;
  a.part_rz = P->part_rz;
  a.part_ftr = P->part_ftr;
  a.u = P->u;
  return a;
}

typedef float f32x4_nt __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4_nt __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }

// y[row] = sum_j A[row][j] xs[j] over one staged chunk, for ROWS rows per wave.
template <int ROWS, typename VecT>
__device__ __forceinline__ void gemv_chunk_f32(const float* __restrict__ A, int64_t lda, int row0, int n, int c0,
                                               int cn, const VecT* xs, int lane, VecT (&acc)[ROWS]) {
  for (int j = lane * 4; j < cn; j += 256) {
    const VecT x0 = xs[j], x1 = j + 1 < cn ? xs[j + 1] : VecT(0), x2 = j + 2 < cn ? xs[j + 2] : VecT(0),
               x3 = j + 3 < cn ? xs[j + 3] : VecT(0);
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const int row = row0 + r;
      if (row < n) {
        // streamed once per product: non-temporal, so x and the CG vectors keep the caches
        const f32x4_nt av = __builtin_nontemporal_load(reinterpret_cast<const f32x4_nt*>(A + (int64_t)row * lda + c0 + j));
        const float4 a = make_float4(av.x, av.y, av.z, av.w);
        VecT s = acc[r];
        s += (VecT)a.x * x0;
        s += (VecT)a.y * x1;
        s += (VecT)a.z * x2;
        s += (VecT)a.w * x3;
        acc[r] = s;
      }
    }
  }
}

template <int ROWS, typename VecT>
__device__ __forceinline__ void gemv_chunk_bf16(const uint16_t* __restrict__ A, int64_t lda, int row0, int n, int c0,
                                                int cn, const VecT* xs, int lane, VecT (&acc)[ROWS]) {
  for (int j = lane * 8; j < cn; j += 512) {
    VecT xv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) xv[q] = j + q < cn ? xs[j + q] : VecT(0);
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const int row = row0 + r;
      if (row < n) {
        const u32x4_nt wv = __builtin_nontemporal_load(reinterpret_cast<const u32x4_nt*>(A + (int64_t)row * lda + c0 + j));
        const uint4 w = make_uint4(wv.x, wv.y, wv.z, wv.w);
        VecT s = acc[r];
        s += (VecT)bf_lo(w.x) * xv[0];
        s += (VecT)bf_hi(w.x) * xv[1];
        s += (VecT)bf_lo(w.y) * xv[2];
        s += (VecT)bf_hi(w.y) * xv[3];
        s += (VecT)bf_lo(w.z) * xv[4];
        s += (VecT)bf_hi(w.z) * xv[5];
        s += (VecT)bf_lo(w.w) * xv[6];
        s += (VecT)bf_hi(w.w) * xv[7];
        acc[r] = s;
      }
    }
  }
}

// Ap = (K + shift I) p for one tile of 4*ROWS rows; part_mv[tile] = sum over the tile of p.Ap
template <int ROWS>
__global__ __launch_bounds__(256) void bcg_matvec(BcgArgs a) {
  __shared__ float xs[kBChunk];
  __shared__ float red[4];
  const int t = blockIdx.x;
  const int s = a.tile_sys[t];
  if (a.done[s]) return;
  const int n = a.n[s];
  const int64_t lda = a.lda[s];
  const int64_t vo = a.voff[s];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row0 = a.tile_row0[t] + wave * ROWS;
  const bool bf = a.prec[s] == 1;
  float acc[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) acc[r] = 0.f;
  for (int c0 = 0; c0 < n; c0 += kBChunk) {
    const int cn = n - c0 < kBChunk ? n - c0 : kBChunk;
    __syncthreads();
    for (int j = threadIdx.x; j < cn; j += 256) xs[j] = a.p[vo + c0 + j];
    __syncthreads();
    if (bf)
      gemv_chunk_bf16<ROWS, float>(a.A16 + a.moff[s], lda, row0, n, c0, cn, xs, lane, acc);
    else
      gemv_chunk_f32<ROWS, float>(a.A32 + a.moff[s], lda, row0, n, c0, cn, xs, lane, acc);
  }
  const float sh = a.shift[s];
  float pap = 0.f;
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const float v = wave_sum(acc[r]);
    const int row = row0 + r;
    if (row < n) {
      const float pv = a.p[vo + row];
      const float y = fmaf(sh, pv, v);
      if (lane == 0) a.Ap[vo + row] = y;
      pap = fmaf(pv, y, pap);
    }
  }
  // pap is identical in every lane of the wave (it uses the broadcast wave sums)
  if (lane == 0) red[wave] = pap;
  __syncthreads();
  if (threadIdx.x == 0) a.part_mv[t] = (red[0] + red[1]) + (red[2] + red[3]);
}

__device__ __forceinline__ double range_sum(const float* part, int b, int e) {
  double acc = 0.0;
  for (int i = b + (int)(threadIdx.x & 63); i < e; i += 64) acc += (double)part[i];
  return wave_sum(acc);
}

__global__ __launch_bounds__(64) void bcg_alpha(BcgArgs a) {
  const int s = blockIdx.x;
  if (a.done[s]) return;
  const double pap = range_sum(a.part_mv, a.mt_begin[s], a.mt_end[s]);
  if (threadIdx.x == 0) {
    const double den = fabs(pap) < 1e-40 ? 1e-40 : pap;
    a.alpha[s] = (float)((double)a.gamma[s] / den);
  }
}

__global__ __launch_bounds__(256) void bcg_xr(BcgArgs a) {
  __shared__ float red[4];
  const int t = blockIdx.x;
  const int s = a.vtile_sys[t];
  if (a.done[s]) return;
  const int n = a.n[s];
  const int64_t vo = a.voff[s];
  const int i0 = a.vtile_i0[t];
  const int i1 = i0 + kBVec < n ? i0 + kBVec : n;
  const float al = a.alpha[s];
  float rr = 0.f;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) {
    const int64_t g = vo + i;
    a.x[g] = fmaf(al, a.p[g], a.x[g]);
    const float rv = fmaf(-al, a.Ap[g], a.r[g]);
    a.r[g] = rv;
    rr = fmaf(rv, rv, rr);
  }
  rr = block_sum<256>(rr, red);
  if (threadIdx.x == 0) a.part_vec[t] = rr;
}

__global__ __launch_bounds__(64) void bcg_beta(BcgArgs a) {
  const int s = blockIdx.x;
  if (a.done[s]) return;
  const double rr = range_sum(a.part_vec, a.vt_begin[s], a.vt_end[s]);
  const double g1 = a.rank > 0 ? range_sum(a.part_rz, a.vt_begin[s], a.vt_end[s]) : rr;
  if (threadIdx.x == 0) {
    const double g0 = (double)a.gamma[s];
    const double den = fabs(g0) < 1e-40 ? 1e-40 : g0;
    a.beta[s] = (float)(g1 / den);
    a.gamma[s] = (float)g1;
    const int k = a.k[s] + 1;
    a.k[s] = k;
    if (sqrt(rr) <= (double)a.thresh[s] || k >= a.max_iters) a.done[s] = 1;
  }
}

__global__ __launch_bounds__(256) void bcg_p(BcgArgs a) {
  const int t = blockIdx.x;
  const int s = a.vtile_sys[t];
  if (a.done[s]) return;
  const int n = a.n[s];
  const int64_t vo = a.voff[s];
  const int i0 = a.vtile_i0[t];
  const int i1 = i0 + kBVec < n ? i0 + kBVec : n;
  const float be = a.beta[s];
  const float* zr = a.rank > 0 ? a.z : a.r;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) {
    const int64_t g = vo + i;
    a.p[g] = fmaf(be, a.p[g], zr[g]);
  }
}

// ------------------------------------------------------------ low-rank preconditioner
// z = (r - F C^-1 F^T r) / shift = (shift I + F F^T)^-1 r per system, in three launches:
// tile partials of F^T r (fp64), per-system u = C^-1 (sum of partials in tile order),
// and the tile update z = (r - F u) / shift with its r.z partial.  F rows are read
// coalesced (lanes over the rank index).
__global__ __launch_bounds__(256) void bcg_pc_ftr(BcgArgs a) {
  __shared__ double red[256];
  const int t = blockIdx.x;
  const int s = a.vtile_sys[t];
  if (a.done[s]) return;
  const int k = a.rank;
  const int groups = 256 / k;
  const int c = threadIdx.x % k, rg = threadIdx.x / k;
  const int n = a.n[s];
  const int64_t vo = a.voff[s];
  const float* Fs = a.F + a.foff[s];
  const int i0 = a.vtile_i0[t];
  const int i1 = i0 + kBVec < n ? i0 + kBVec : n;
  // four independent partial sums per thread keep several row loads in flight
  double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
  if (rg < groups) {
    int i = i0 + rg;
    for (; i + 3 * groups < i1; i += 4 * groups) {
      acc0 = fma((double)Fs[(int64_t)i * k + c], (double)a.r[vo + i], acc0);
      acc1 = fma((double)Fs[(int64_t)(i + groups) * k + c], (double)a.r[vo + i + groups], acc1);
      acc2 = fma((double)Fs[(int64_t)(i + 2 * groups) * k + c], (double)a.r[vo + i + 2 * groups], acc2);
      acc3 = fma((double)Fs[(int64_t)(i + 3 * groups) * k + c], (double)a.r[vo + i + 3 * groups], acc3);
    }
    for (; i < i1; i += groups) acc0 = fma((double)Fs[(int64_t)i * k + c], (double)a.r[vo + i], acc0);
  }
  red[threadIdx.x] = (acc0 + acc1) + (acc2 + acc3);
  __syncthreads();
  if (threadIdx.x < k) {
    double sum = 0.0;
    for (int g = 0; g < groups; ++g) sum += red[g * k + threadIdx.x];
    a.part_ftr[(int64_t)t * k + threadIdx.x] = sum;
  }
}

__global__ __launch_bounds__(64) void bcg_pc_core(BcgArgs a) {
  __shared__ double ts[64];
  const int s = blockIdx.x;
  if (a.done[s]) return;
  const int k = a.rank;
  const int c = threadIdx.x;
  if (c < k) {
    double acc = 0.0;
    for (int t = a.vt_begin[s]; t < a.vt_end[s]; ++t) acc += a.part_ftr[(int64_t)t * k + c];
    ts[c] = acc;
  }
  __syncthreads();
  if (c < k) {
    const double* Ci = a.Cinv + (int64_t)s * k * k + (int64_t)c * k;
    double acc = 0.0;
    for (int j = 0; j < k; ++j) acc = fma(Ci[j], ts[j], acc);
    a.u[(int64_t)s * k + c] = acc;
  }
}

__global__ __launch_bounds__(256) void bcg_pc_apply(BcgArgs a) {
  __shared__ double us[64];
  __shared__ float red[4];
  const int t = blockIdx.x;
  const int s = a.vtile_sys[t];
  if (a.done[s]) return;
  const int k = a.rank;
  for (int c = threadIdx.x; c < k; c += 256) us[c] = a.u[(int64_t)s * k + c];
  __syncthreads();
  const int n = a.n[s];
  const int64_t vo = a.voff[s];
  const float* Fs = a.F + a.foff[s];
  const int i0 = a.vtile_i0[t];
  const int i1 = i0 + kBVec < n ? i0 + kBVec : n;
  const double inv = 1.0 / (double)a.shift[s];
  float rz = 0.f;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) {
    const float* fr = Fs + (int64_t)i * k;
    double acc = 0.0;
    for (int c = 0; c < k; ++c) acc = fma((double)fr[c], us[c], acc);
    const double rv = (double)a.r[vo + i];
    const float zv = (float)((rv - acc) * inv);
    a.z[vo + i] = zv;
    rz = fmaf((float)rv, zv, rz);
  }
  rz = block_sum<256>(rz, red);
  if (threadIdx.x == 0) a.part_rz[t] = rz;
}

void enqueue_precond(const ck_bcg_plan* P, const BcgArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(bcg_pc_ftr, dim3(P->ntiles_vec), dim3(256), 0, s, a);
  hipLaunchKernelGGL(bcg_pc_core, dim3(P->nsys), dim3(64), 0, s, a);
  hipLaunchKernelGGL(bcg_pc_apply, dim3(P->ntiles_vec), dim3(256), 0, s, a);
}

// start, preconditioned: p = z for active systems
__global__ __launch_bounds__(256) void bcg_start_p(BcgArgs a) {
  const int t = blockIdx.x;
  const int s = a.vtile_sys[t];
  if (a.done[s]) return;
  const int n = a.n[s];
  const int64_t vo = a.voff[s];
  const int i0 = a.vtile_i0[t];
  const int i1 = i0 + kBVec < n ? i0 + kBVec : n;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) a.p[vo + i] = a.z[vo + i];
}

// start: x = 0, r = p = b for active systems (done == 0 on entry), tile ||b||^2
__global__ __launch_bounds__(256) void bcg_start_vec(BcgArgs a, const float* __restrict__ b) {
  __shared__ float red[4];
  const int t = blockIdx.x;
  const int s = a.vtile_sys[t];
  if (a.done[s]) return;
  const int n = a.n[s];
  const int64_t vo = a.voff[s];
  const int i0 = a.vtile_i0[t];
  const int i1 = i0 + kBVec < n ? i0 + kBVec : n;
  float bb = 0.f;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) {
    const int64_t g = vo + i;
    const float bv = b[g];
    a.x[g] = 0.f;
    a.r[g] = bv;
    a.p[g] = bv;
    bb = fmaf(bv, bv, bb);
  }
  bb = block_sum<256>(bb, red);
  if (threadIdx.x == 0) a.part_vec[t] = bb;
}

__global__ __launch_bounds__(64) void bcg_start_sys(BcgArgs a, float tol_rel, float tol_abs) {
  const int s = blockIdx.x;
  if (a.done[s]) return;
  const double g = range_sum(a.part_vec, a.vt_begin[s], a.vt_end[s]);
  const double gz = a.rank > 0 ? range_sum(a.part_rz, a.vt_begin[s], a.vt_end[s]) : g;
  if (threadIdx.x == 0) {
    a.gamma[s] = (float)gz;
    const float th = tol_rel * (float)sqrt(g) + tol_abs;
    a.thresh[s] = th;
    a.k[s] = 0;
    if (sqrt(g) <= (double)th) a.done[s] = 1;
  }
}

// ------------------------------------------------------------ FP64 refinement
// r = b - (K + shift I) x with fp32 K as stored and fp64 x, b, accumulation; r32 = (float) r;
// part_mv gets the tile's r.r and part2 the tile's b.b (both fp64).  Systems with
// active[s] == 0 are skipped.
template <int ROWS>
__global__ __launch_bounds__(256) void bcg_residual_f64(BcgArgs a, const int32_t* __restrict__ active,
                                                        const double* __restrict__ x, const double* __restrict__ b,
                                                        float* __restrict__ r32, double* __restrict__ prr,
                                                        double* __restrict__ pbb) {
  __shared__ double xs[kBChunk];
  __shared__ double red[8];
  const int t = blockIdx.x;
  const int s = a.tile_sys[t];
  if (!active[s]) return;
  const int n = a.n[s];
  const int64_t lda = a.lda[s];
  const int64_t vo = a.voff[s];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row0 = a.tile_row0[t] + wave * ROWS;
  double acc[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) acc[r] = 0.0;
  for (int c0 = 0; c0 < n; c0 += kBChunk) {
    const int cn = n - c0 < kBChunk ? n - c0 : kBChunk;
    __syncthreads();
    for (int j = threadIdx.x; j < cn; j += 256) xs[j] = x[vo + c0 + j];
    __syncthreads();
    gemv_chunk_f32<ROWS, double>(a.A32 + a.moff[s], lda, row0, n, c0, cn, xs, lane, acc);
  }
  const double sh = (double)a.shift[s];
  double rr = 0.0, bb = 0.0;
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const double v = wave_sum(acc[r]);
    const int row = row0 + r;
    if (row < n) {
      const double bv = b[vo + row];
      const double rv = bv - (v + sh * x[vo + row]);
      if (lane == 0) r32[vo + row] = (float)rv;
      rr = fma(rv, rv, rr);
      bb = fma(bv, bv, bb);
    }
  }
  if (lane == 0) {
    red[wave] = rr;
    red[4 + wave] = bb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    prr[t] = (red[0] + red[1]) + (red[2] + red[3]);
    pbb[t] = (red[4] + red[5]) + (red[6] + red[7]);
  }
}

__global__ __launch_bounds__(64) void bcg_norms_f64(BcgArgs a, const int32_t* __restrict__ active,
                                                    const double* __restrict__ prr, const double* __restrict__ pbb,
                                                    double* __restrict__ rel) {
  const int s = blockIdx.x;
  if (!active[s]) return;
  double rr = 0.0, bb = 0.0;
  for (int i = a.mt_begin[s] + (int)threadIdx.x; i < a.mt_end[s]; i += 64) {
    rr += prr[i];
    bb += pbb[i];
  }
  rr = wave_sum(rr);
  bb = wave_sum(bb);
  if (threadIdx.x == 0) rel[s] = sqrt(rr) / (bb > 0.0 ? sqrt(bb) : 1.0);
}

// x64 += x32 on active systems
__global__ __launch_bounds__(256) void bcg_update_f64(BcgArgs a, const int32_t* __restrict__ active,
                                                      double* __restrict__ x64) {
  const int t = blockIdx.x;
  const int s = a.vtile_sys[t];
  if (!active[s]) return;
  const int n = a.n[s];
  const int64_t vo = a.voff[s];
  const int i0 = a.vtile_i0[t];
  const int i1 = i0 + kBVec < n ? i0 + kBVec : n;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) x64[vo + i] += (double)a.x[vo + i];
}

int check_plan(const ck_bcg_plan* P) {
  if (!P || P->nsys < 1 || P->ntiles_mv < 1 || P->ntiles_vec < 1 || !P->tile_sys || !P->tile_row0 ||
      !P->vtile_sys || !P->vtile_i0 || !P->n || !P->lda || !P->voff || !P->moff || !P->mt_begin || !P->mt_end ||
      !P->vt_begin || !P->vt_end || !P->shift || !P->prec || !P->A32 || !P->x || !P->r || !P->p || !P->Ap ||
      !P->part_mv || !P->part_vec || !P->gamma || !P->alpha || !P->beta || !P->thresh || !P->k || !P->done)
    return CK_ERR_BAD_ARG;
  const int R = P->rows_per_tile;
  if (R != 4 && R != 8 && R != 16 && R != 32) return CK_ERR_UNSUPPORTED;
  if (P->rank < 0 || P->rank > 64 || (P->rank > 0 && 256 % P->rank)) return CK_ERR_UNSUPPORTED;
  if (P->rank > 0 && (!P->F || !P->foff || !P->Cinv || !P->z || !P->part_rz || !P->part_ftr || !P->u))
    return CK_ERR_BAD_ARG;
  return CK_OK;
}

template <int ROWS>
void launch_matvec(const BcgArgs& a, int ntiles, hipStream_t s) {
  hipLaunchKernelGGL(bcg_matvec<ROWS>, dim3(ntiles), dim3(256), 0, s, a);
}

void matvec_dispatch(const ck_bcg_plan* P, const BcgArgs& a, hipStream_t s) {
  switch (P->rows_per_tile) {
    case 4: launch_matvec<1>(a, P->ntiles_mv, s); break;
    case 8: launch_matvec<2>(a, P->ntiles_mv, s); break;
    case 16: launch_matvec<4>(a, P->ntiles_mv, s); break;
    default: launch_matvec<8>(a, P->ntiles_mv, s); break;
  }
}

void enqueue_iterations(const ck_bcg_plan* P, int iters, hipStream_t s) {
  const BcgArgs a = to_args(P);
  for (int it = 0; it < iters; ++it) {
    matvec_dispatch(P, a, s);
    hipLaunchKernelGGL(bcg_alpha, dim3(P->nsys), dim3(64), 0, s, a);
    hipLaunchKernelGGL(bcg_xr, dim3(P->ntiles_vec), dim3(256), 0, s, a);
    if (P->rank > 0) enqueue_precond(P, a, s);
    hipLaunchKernelGGL(bcg_beta, dim3(P->nsys), dim3(64), 0, s, a);
    hipLaunchKernelGGL(bcg_p, dim3(P->ntiles_vec), dim3(256), 0, s, a);
  }
}

struct CkGraph {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

}  // namespace

extern "C" {

int ck_bcg_start(const ck_bcg_plan* plan, const float* b, float tol_rel, float tol_abs, void* stream) {
  const int st = check_plan(plan);
  if (st != CK_OK) return st;
  if (!b) return CK_ERR_BAD_ARG;
  hipStream_t s = (hipStream_t)stream;
  const BcgArgs a = to_args(plan);
  hipLaunchKernelGGL(bcg_start_vec, dim3(plan->ntiles_vec), dim3(256), 0, s, a, b);
  if (plan->rank > 0) enqueue_precond(plan, a, s);
  // the threshold test in bcg_start_sys may mark systems done: p = z runs first
  if (plan->rank > 0) hipLaunchKernelGGL(bcg_start_p, dim3(plan->ntiles_vec), dim3(256), 0, s, a);
  hipLaunchKernelGGL(bcg_start_sys, dim3(plan->nsys), dim3(64), 0, s, a, tol_rel, tol_abs);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_bcg_iterate(const ck_bcg_plan* plan, int32_t iters, void* stream) {
  const int st = check_plan(plan);
  if (st != CK_OK) return st;
  if (iters < 1) return CK_ERR_BAD_ARG;
  enqueue_iterations(plan, iters, (hipStream_t)stream);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_bcg_matvec(const ck_bcg_plan* plan, void* stream) {
  const int st = check_plan(plan);
  if (st != CK_OK) return st;
  matvec_dispatch(plan, to_args(plan), (hipStream_t)stream);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_bcg_graph_create(const ck_bcg_plan* plan, int32_t iters, void* stream, void** graph_out) {
  const int st = check_plan(plan);
  if (st != CK_OK) return st;
  if (iters < 1 || !graph_out || !stream) return CK_ERR_BAD_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return CK_ERR_HIP;
  enqueue_iterations(plan, iters, s);
  const hipError_t launch_err = hipGetLastError();
  hipGraph_t graph = nullptr;
  const hipError_t end_err = hipStreamEndCapture(s, &graph);
  if (launch_err != hipSuccess || end_err != hipSuccess || !graph) {
    if (graph) hipGraphDestroy(graph);
    return CK_ERR_HIP;
  }
  hipGraphExec_t exec = nullptr;
  if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    hipGraphDestroy(graph);
    return CK_ERR_HIP;
  }
  *graph_out = new CkGraph{graph, exec};
  return CK_OK;
}

int ck_graph_launch(void* graph, void* stream) {
  if (!graph) return CK_ERR_BAD_ARG;
  if (hipGraphLaunch(static_cast<CkGraph*>(graph)->exec, (hipStream_t)stream) != hipSuccess) return CK_ERR_HIP;
  return CK_OK;
}

int ck_graph_destroy(void* graph) {
  if (!graph) return CK_ERR_BAD_ARG;
  CkGraph* g = static_cast<CkGraph*>(graph);
  hipGraphExecDestroy(g->exec);
  hipGraphDestroy(g->graph);
  delete g;
  return CK_OK;
}

int ck_bcg_residual_f64(const ck_bcg_plan* plan, const int32_t* active, const double* x64, const double* b64,
                        float* r32, double* part_rr, double* part_bb, double* rel, void* stream) {
  const int st = check_plan(plan);
  if (st != CK_OK) return st;
  if (!active || !x64 || !b64 || !r32 || !part_rr || !part_bb || !rel) return CK_ERR_BAD_ARG;
  hipStream_t s = (hipStream_t)stream;
  const BcgArgs a = to_args(plan);
  switch (plan->rows_per_tile) {
    case 4:
      hipLaunchKernelGGL(bcg_residual_f64<1>, dim3(plan->ntiles_mv), dim3(256), 0, s, a, active, x64, b64, r32,
                         part_rr, part_bb);
      break;
    case 8:
      hipLaunchKernelGGL(bcg_residual_f64<2>, dim3(plan->ntiles_mv), dim3(256), 0, s, a, active, x64, b64, r32,
                         part_rr, part_bb);
      break;
    case 16:
      hipLaunchKernelGGL(bcg_residual_f64<4>, dim3(plan->ntiles_mv), dim3(256), 0, s, a, active, x64, b64, r32,
                         part_rr, part_bb);
      break;
    default:
      hipLaunchKernelGGL(bcg_residual_f64<8>, dim3(plan->ntiles_mv), dim3(256), 0, s, a, active, x64, b64, r32,
                         part_rr, part_bb);
      break;
  }
  hipLaunchKernelGGL(bcg_norms_f64, dim3(plan->nsys), dim3(64), 0, s, a, active, (const double*)part_rr,
                     (const double*)part_bb, rel);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_bcg_update_f64(const ck_bcg_plan* plan, const int32_t* active, double* x64, void* stream) {
  const int st = check_plan(plan);
  if (st != CK_OK) return st;
  if (!active || !x64) return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(bcg_update_f64, dim3(plan->ntiles_vec), dim3(256), 0, (hipStream_t)stream, to_args(plan),
                     active, x64);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int32_t ck_bcg_vec_tile(void) { return kBVec; }

}  // extern "C"
