// Probe streams, Hutchinson accumulation, the quadrature reduction of
// stochastic Lanczos quadrature, and Ritz-vector assembly.
//
// CSR Hutchinson with packed probes: 64 Rademacher probes of a row live in one
// 64-bit word, so a wavefront takes one row with lane p = probe p; the row's
// values and column indices are wave-uniform loads and each nonzero costs one
// 8-byte gather of the column's word.  The trace-only variant never forms a
// per-probe value at all: sum_p z_p[i] z_p[j] = 64 - 2 popcount(w_i ^ w_j).
#include "common.hpp"
#include "ops_internal.hpp"

using namespace ck;

namespace {

// ---------------------------------------------------------------- probes
__global__ __launch_bounds__(256) void rng_rademacher_kernel(int64_t n, int nprobes, uint64_t seed,
                                                             float* __restrict__ Z, int64_t ldz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int g = blockIdx.y;
  if (i >= n) return;
  const uint64_t w = stream_word(stream_key(seed, (uint64_t)g), (uint64_t)i);
  const int p0 = 64 * g;
  const int p1 = p0 + 64 < nprobes ? p0 + 64 : nprobes;
  for (int p = p0; p < p1; ++p) Z[(int64_t)p * ldz + i] = ((w >> (p - p0)) & 1ull) ? -1.f : 1.f;
}

__global__ __launch_bounds__(256) void rng_bits_kernel(int64_t n, uint64_t seed, uint64_t* __restrict__ W) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int g = blockIdx.y;
  if (i >= n) return;
  W[(int64_t)g * n + i] = stream_word(stream_key(seed, (uint64_t)g), (uint64_t)i);
}

__global__ __launch_bounds__(256) void rng_gaussian_kernel(int64_t n, uint64_t seed, float* __restrict__ Z,
                                                           int64_t ldz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (i >= n) return;
  const uint64_t w = stream_word(stream_key(seed ^ kGaussSalt, (uint64_t)p), (uint64_t)i);
  Z[(int64_t)p * ldz + i] = gauss_from_word(w);
}

// ---------------------------------------------------------------- Hutchinson
__global__ __launch_bounds__(256) void hutch_accum_kernel(int64_t n, int nprobes, const float* __restrict__ Z,
                                                          int64_t ldz, const float* __restrict__ AZ, int64_t ldaz,
                                                          float* __restrict__ sum, float* __restrict__ sumsq) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f, s2 = 0.f;
  for (int p = 0; p < nprobes; ++p) {
    const float est = AZ[(int64_t)p * ldaz + i] * Z[(int64_t)p * ldz + i];
    s += est;
    s2 = fmaf(est, est, s2);
  }
  sum[i] += s;
  sumsq[i] += s2;
}

__device__ __forceinline__ float sign_bit(uint64_t w, int lane) { return ((w >> lane) & 1ull) ? -1.f : 1.f; }

__global__ __launch_bounds__(256) void hutch_csr_bits_kernel(int n, const int32_t* __restrict__ rowptr,
                                                             const int32_t* __restrict__ colind,
                                                             const float* __restrict__ val, float scale, float shift,
                                                             int ngroups, const uint64_t* __restrict__ W,
                                                             float* __restrict__ sum, float* __restrict__ sumsq) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (row >= n) return;
  const int k0 = rowptr[row], k1 = rowptr[row + 1];
  float s = 0.f, s2 = 0.f;
  for (int g = 0; g < ngroups; ++g) {
    const uint64_t* Wg = W + (int64_t)g * n;
    const float zi = sign_bit(Wg[row], lane);
    float acc = 0.f;
    for (int k = k0; k < k1; ++k) acc = fmaf(val[k], sign_bit(Wg[colind[k]], lane), acc);
    const float az = fmaf(shift, zi, scale * acc);
    const float est = zi * az;
    s += wave_sum(est);
    s2 += wave_sum(est * est);
  }
  if (lane == 0) {
    sum[row] += s;
    sumsq[row] += s2;
  }
}

constexpr int kBitsShort = 8;
constexpr int kTraceRowsPerThread = 4;  // rows b*1024 + t + 256 j, j < 4: four rows' loads in flight

// Probe words of the trace pass: stored (W from ck_rng_rademacher_bits) or recomputed from
// the seed.  Both give stream_word(stream_key(seed, g), i) for word (g, i).
struct StoredWords {
  const uint64_t* W;
  int64_t n;
  __device__ uint64_t operator()(int g, int64_t i) const { return W[(int64_t)g * n + i]; }
};
struct SeededWords {
  uint64_t seed;
  __device__ uint64_t operator()(int g, int64_t i) const {
    return stream_word(stream_key(seed, (uint64_t)g), (uint64_t)i);
  }
};

// sum_k val_k * sum_g (64 - 2 popcount(w_g(row) ^ w_g(col_k))) for one row, k ascending.
// Short rows with one word group load their column indices and values in one batch.
template <typename Words>
__device__ __forceinline__ double trace_row(int row, const int32_t* __restrict__ rowptr,
                                            const int32_t* __restrict__ colind, const float* __restrict__ val,
                                            int ngroups, const Words& words) {
  const int k0 = rowptr[row], k1 = rowptr[row + 1];
  const int len = k1 - k0;
  double acc = 0.0;
  if (ngroups == 1 && len <= kBitsShort) {
    const uint64_t wr = words(0, row);
    int c[kBitsShort];
    float v[kBitsShort];
#pragma unroll
    for (int i = 0; i < kBitsShort; ++i) {
      c[i] = i < len ? colind[k0 + i] : row;
      v[i] = i < len ? val[k0 + i] : 0.f;
    }
    uint64_t wc[kBitsShort];
#pragma unroll
    for (int i = 0; i < kBitsShort; ++i) wc[i] = i < len ? words(0, c[i]) : wr;
    // __popcll returns unsigned: convert before the subtraction can wrap
#pragma unroll
    for (int i = 0; i < kBitsShort; ++i)
      if (i < len) acc += (double)v[i] * (double)(64 - 2 * (int)__popcll(wr ^ wc[i]));
  } else {
    for (int k = k0; k < k1; ++k) {
      const int c = colind[k];
      int agree = 0;
      for (int g = 0; g < ngroups; ++g) agree += 64 - 2 * (int)__popcll(words(g, row) ^ words(g, c));
      acc += (double)val[k] * (double)agree;
    }
  }
  return acc;
}

// One partial per 1024 rows: a thread adds its four rows' sums (j ascending), then the block
// sum (fixed order), so the stored-word and seeded passes give the same bits.  With one
// word group and four short rows (the common case) the four rows' extents are loaded
// together and then all their entries in one batch (indexes clamped into [0, nnz), the
// values selected afterwards), so a thread waits for two loads rather than eight; each
// row's sum is formed as in trace_row.
template <typename Words>
__global__ __launch_bounds__(256) void trace_partials(int n, int64_t nnz, const int32_t* __restrict__ rowptr,
                                                      const int32_t* __restrict__ colind,
                                                      const float* __restrict__ val, int ngroups, Words words,
                                                      double* __restrict__ part) {
  __shared__ double red[4];
  constexpr int R = kTraceRowsPerThread;
  double acc = 0.0;
  int row[R], k0[R], len[R];
  bool all_short = ngroups == 1;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    row[j] = blockIdx.x * (256 * R) + j * 256 + threadIdx.x;
    const int rc = row[j] < n ? row[j] : n - 1;
    k0[j] = rowptr[rc];
    len[j] = rowptr[rc + 1] - k0[j];
    all_short = all_short && (row[j] >= n || len[j] <= kBitsShort);
  }
  if (all_short && nnz > 0) {
    const int64_t last = nnz - 1;
    int c[R][kBitsShort];
    float v[R][kBitsShort];
#pragma unroll
    for (int j = 0; j < R; ++j)
#pragma unroll
      for (int i = 0; i < kBitsShort; ++i) {
        const int64_t k = (int64_t)k0[j] + i < last ? (int64_t)k0[j] + i : last;
        c[j][i] = colind[k];
        v[j][i] = val[k];
      }
#pragma unroll
    for (int j = 0; j < R; ++j) {
      if (row[j] >= n) continue;
      const uint64_t wr = words(0, row[j]);
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < kBitsShort; ++i)
        if (i < len[j]) s += (double)v[j][i] * (double)(64 - 2 * (int)__popcll(wr ^ words(0, c[j][i])));
      acc += s;
    }
  } else {
#pragma unroll
    for (int j = 0; j < R; ++j)
      if (row[j] < n) acc += trace_row(row[j], rowptr, colind, val, ngroups, words);
  }
  acc = block_sum<256>(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// The one-word-group pass with coalesced entry loads: each wave takes 64 consecutive rows,
// loads their nonzeros [rowptr[r], rowptr[r + 64]) as whole 64-entry lines (column index and
// value, up to kStreamStage entries) into its own LDS stage, and each lane then walks its
// row's entries from LDS; rows of a chunk with more entries take trace_row's loads.  The
// block's kStreamRows rows have their probe words in LDS (one mixed word per row), a column
// outside them is mixed again or gathered.  Each row's entries are summed in order, rows in
// round order, then the block sum: one partial per kStreamRows rows.
constexpr int kStreamRounds = 2;
constexpr int kStreamRows = 256 * kStreamRounds;
constexpr int kStreamStage = 512;

template <typename Words>
__global__ __launch_bounds__(256) void trace_stream(int n, const int32_t* __restrict__ rowptr,
                                                    const int32_t* __restrict__ colind,
                                                    const float* __restrict__ val, Words words,
                                                    double* __restrict__ part) {
  __shared__ uint64_t wl[kStreamRows];
  __shared__ int sc[4][kStreamStage];
  __shared__ float sv[4][kStreamStage];
  __shared__ double red[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r0 = blockIdx.x * kStreamRows;
#pragma unroll
  for (int j = 0; j < kStreamRounds; ++j) {
    const int row = r0 + j * 256 + threadIdx.x;
    wl[j * 256 + threadIdx.x] = row < n ? words(0, row) : 0ull;
  }
  __syncthreads();
  auto word = [&](int c) -> uint64_t {
    const int d = c - r0;
    return (d >= 0 && d < kStreamRows) ? wl[d] : words(0, c);
  };
  double acc = 0.0;
  for (int round = 0; round < kStreamRounds; ++round) {
    const int base = r0 + round * 256 + wave * 64;
    if (base >= n) break;  // wave-uniform: no block barrier follows
    const int row = base + lane;
    const int last = min(base + 63, n - 1);
    const int rc = min(row, last);
    const int rs = rowptr[rc], re = rowptr[rc + 1];
    const int p0 = __shfl(rs, 0), p1 = __shfl(re, last - base);
    const int cnt = p1 - p0;
    const int len = row <= last ? re - rs : 0;
    const uint64_t wr = wl[row - r0];
    double rsum = 0.0;
    if (cnt <= kStreamStage) {
      if (cnt > 0) {
        // every line's loads first (indexes clamped into the chunk), then the LDS writes
        int cv[kStreamStage / 64];
        float vv[kStreamStage / 64];
#pragma unroll
        for (int t = 0; t < kStreamStage / 64; ++t) {
          const int k = min(p0 + t * 64 + lane, p1 - 1);
          cv[t] = colind[k];
          vv[t] = val[k];
        }
#pragma unroll
        for (int t = 0; t < kStreamStage / 64; ++t) {
          sc[wave][t * 64 + lane] = cv[t];
          sv[wave][t * 64 + lane] = vv[t];
        }
      }
      // the stage is this wave's: order its LDS writes before the lanes' reads
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int e0 = rs - p0;
      for (int e = 0; e < len; ++e) {
        const int c = sc[wave][e0 + e];
        const float v = sv[wave][e0 + e];
        rsum += (double)v * (double)(64 - 2 * (int)__popcll(wr ^ word(c)));
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
      for (int k = rs; k < rs + len; ++k) {
        const int c = colind[k];
        rsum += (double)val[k] * (double)(64 - 2 * (int)__popcll(wr ^ word(c)));
      }
    }
    acc += rsum;
  }
  acc = block_sum<256>(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void trace_bits_final(int nparts, const double* __restrict__ part, double scale,
                                                        double shift_n, double inv_probes, double* __restrict__ out) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) acc += part[i];
  acc = block_sum<256>(acc, red);
  if (threadIdx.x == 0) out[0] = scale * acc * inv_probes + shift_n;
}

// ---------------------------------------------------------------- quadrature
__device__ double apply_f(int kind, double x) {
  switch (kind) {
    case 0:
      return log(x);
    case 1:
      return x;
    case 2:
      return exp(x);
    case 3:
      return sqrt(x);
    default:
      return 1.0 / x;
  }
}

__global__ __launch_bounds__(64) void slq_kernel(int m, const double* __restrict__ evals,
                                                 const double* __restrict__ first_row,
                                                 const double* __restrict__ scale, int f_kind, double eps,
                                                 int max_mode, double* __restrict__ out) {
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  double mx = -INFINITY;
  for (int c = lane; c < m; c += 64) {
    const double t = evals[(int64_t)b * m + c];
    mx = fmax(mx, max_mode == 0 ? fabs(t) : t);
  }
  // wave max in fp64 via two passes of readlane (m is small)
  for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  const double thresh = 10.0 * eps * mx;
  double acc = 0.0;
  for (int c = lane; c < m; c += 64) {
    const double t = evals[(int64_t)b * m + c];
    const double tau = first_row[(int64_t)b * m + c];
    const double fv = fabs(t) > thresh ? apply_f(f_kind, t) : 0.0;
    acc += tau * tau * fv;
  }
  acc = wave_sum(acc);
  if (lane == 0) out[b] = (scale ? scale[b] : 1.0) * acc;
}

// ---------------------------------------------------------------- Ritz vectors
// out = S^T V for up to 32 columns of S per block: S is staged through LDS as fp32 in slices
// of 128 basis rows, every thread keeps its row's 32 sums in registers, so V is read once per
// 32 Ritz vectors and each loaded element feeds 32 FMAs.
constexpr int kCombine = 32;
constexpr int kCombineRows = 128;

__global__ __launch_bounds__(256) void basis_combine_kernel(int64_t n, int m, int k, const float* __restrict__ V,
                                                            int64_t ldv, const double* __restrict__ S, int64_t lds_,
                                                            float* __restrict__ out, int64_t ldo) {
  __shared__ float Ss[kCombineRows][kCombine];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c0 = blockIdx.y * kCombine;
  const int64_t ic = i < n ? i : n - 1;
  float acc[kCombine];
#pragma unroll
  for (int c = 0; c < kCombine; ++c) acc[c] = 0.f;
  for (int j0 = 0; j0 < m; j0 += kCombineRows) {
    const int rows = min(kCombineRows, m - j0);
    __syncthreads();
    for (int t = threadIdx.x; t < kCombineRows * kCombine; t += blockDim.x) {
      const int r = t / kCombine, c = t % kCombine;
      Ss[r][c] = (r < rows && c0 + c < k) ? (float)S[(int64_t)(j0 + r) * lds_ + c0 + c] : 0.f;
    }
    __syncthreads();
    const float* Vj = V + (int64_t)j0 * ldv + ic;
#pragma unroll 4
    for (int r = 0; r < rows; ++r) {
      const float v = Vj[(int64_t)r * ldv];
#pragma unroll
      for (int c = 0; c < kCombine; ++c) acc[c] = fmaf(Ss[r][c], v, acc[c]);
    }
  }
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < kCombine; ++c)
    if (c0 + c < k) out[(int64_t)(c0 + c) * ldo + i] = acc[c];
}

// out[b][i] = sum_j C[b][j] V[b][j][i]: one combination per probe of its own basis.
__global__ __launch_bounds__(256) void basis_combine_batched_kernel(int64_t n, int k, const float* __restrict__ V,
                                                                    int64_t ldv, int64_t ldb,
                                                                    const double* __restrict__ C, int64_t ldc,
                                                                    float* __restrict__ out, int64_t ldo) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (i >= n) return;
  const float* Vb = V + (int64_t)b * ldb;
  float acc = 0.f;
  for (int j = 0; j < k; ++j) acc = fmaf((float)C[(int64_t)b * ldc + j], Vb[(int64_t)j * ldv + i], acc);
  out[(int64_t)b * ldo + i] = acc;
}

}  // namespace

extern "C" {

int ck_basis_combine_batched(int64_t n, int32_t nb, int32_t k, const float* V, int64_t ldv, int64_t ldb,
                             const double* C, int64_t ldc, float* out, int64_t ldo, void* stream) {
  if (n < 1 || nb < 1 || k < 1 || !V || !C || !out || ldv < n || ldo < n || ldc < k || ldb < (int64_t)k * ldv)
    return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(basis_combine_batched_kernel, dim3((unsigned)cdiv(n, 256), nb), dim3(256), 0,
                     (hipStream_t)stream, n, k, V, ldv, ldb, C, ldc, out, ldo);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_rng_rademacher_f32(int64_t n, int32_t nprobes, uint64_t seed, float* Z, int64_t ldz, void* stream) {
  if (n < 1 || nprobes < 1 || !Z || ldz < n) return CK_ERR_BAD_ARG;
  const int groups = (nprobes + 63) / 64;
  hipLaunchKernelGGL(rng_rademacher_kernel, dim3((unsigned)cdiv(n, 256), groups), dim3(256), 0, (hipStream_t)stream, n,
                     nprobes, seed, Z, ldz);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_rng_gaussian_f32(int64_t n, int32_t nprobes, uint64_t seed, float* Z, int64_t ldz, void* stream) {
  if (n < 1 || nprobes < 1 || !Z || ldz < n) return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(rng_gaussian_kernel, dim3((unsigned)cdiv(n, 256), nprobes), dim3(256), 0, (hipStream_t)stream, n,
                     seed, Z, ldz);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_rng_rademacher_bits(int64_t n, int32_t ngroups, uint64_t seed, uint64_t* W, void* stream) {
  if (n < 1 || ngroups < 1 || !W) return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(rng_bits_kernel, dim3((unsigned)cdiv(n, 256), ngroups), dim3(256), 0, (hipStream_t)stream, n, seed,
                     W);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_hutch_accumulate(int64_t n, int32_t nprobes, const float* Z, int64_t ldz, const float* AZ, int64_t ldaz,
                        float* sum, float* sumsq, void* stream) {
  if (n < 1 || nprobes < 1 || !Z || !AZ || !sum || !sumsq || ldz < n || ldaz < n) return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(hutch_accum_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, n, nprobes,
                     Z, ldz, AZ, ldaz, sum, sumsq);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_hutch_csr_bits(const ck_op* A, int32_t ngroups, const uint64_t* W, float* sum, float* sumsq, void* stream) {
  if (!A || A->kind != CK_OP_CSR_F32 || !A->rowptr || !A->colind || !A->values || ngroups < 1 || !W || !sum || !sumsq)
    return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(hutch_csr_bits_kernel, dim3((unsigned)cdiv(A->n, 4)), dim3(256), 0, (hipStream_t)stream, A->n,
                     A->rowptr, A->colind, A->values, A->scale, A->shift, ngroups, W, sum, sumsq);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

size_t ck_trace_workspace_bytes(int64_t n) { return sizeof(double) * (size_t)cdiv(n, 256); }

int ck_trace_csr_rademacher_bits(const ck_op* A, int32_t ngroups, const uint64_t* W, double* out, void* work,
                                 size_t work_bytes, void* stream) {
  if (!A || A->kind != CK_OP_CSR_F32 || !A->rowptr || !A->colind || !A->values || ngroups < 1 || !W || !out)
    return CK_ERR_BAD_ARG;
  if (!work || work_bytes < ck_trace_workspace_bytes(A->n)) return CK_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  const bool stream1 = ngroups == 1;
  const int nparts = (int)(stream1 ? cdiv(A->n, kStreamRows) : cdiv(A->n, 256 * kTraceRowsPerThread));
  double* part = static_cast<double*>(work);
  if (stream1)
    hipLaunchKernelGGL(trace_stream<StoredWords>, dim3(nparts), dim3(256), 0, s, A->n, A->rowptr, A->colind, A->values,
                       StoredWords{W, (int64_t)A->n}, part);
  else
    hipLaunchKernelGGL(trace_partials<StoredWords>, dim3(nparts), dim3(256), 0, s, A->n, A->nnz, A->rowptr, A->colind,
                       A->values, ngroups, StoredWords{W, (int64_t)A->n}, part);
  hipLaunchKernelGGL(trace_bits_final, dim3(1), dim3(256), 0, s, nparts, (const double*)part, (double)A->scale,
                     (double)A->shift * (double)A->n, 1.0 / (64.0 * ngroups), out);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_trace_csr_rademacher_seeded(const ck_op* A, int32_t ngroups, uint64_t seed, double* out, void* work,
                                   size_t work_bytes, void* stream) {
  if (!A || A->kind != CK_OP_CSR_F32 || !A->rowptr || !A->colind || !A->values || ngroups < 1 || !out)
    return CK_ERR_BAD_ARG;
  if (!work || work_bytes < ck_trace_workspace_bytes(A->n)) return CK_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  const bool stream1 = ngroups == 1;
  const int nparts = (int)(stream1 ? cdiv(A->n, kStreamRows) : cdiv(A->n, 256 * kTraceRowsPerThread));
  double* part = static_cast<double*>(work);
  if (stream1)
    hipLaunchKernelGGL(trace_stream<SeededWords>, dim3(nparts), dim3(256), 0, s, A->n, A->rowptr, A->colind, A->values,
                       SeededWords{seed}, part);
  else
    hipLaunchKernelGGL(trace_partials<SeededWords>, dim3(nparts), dim3(256), 0, s, A->n, A->nnz, A->rowptr, A->colind,
                       A->values, ngroups, SeededWords{seed}, part);
  hipLaunchKernelGGL(trace_bits_final, dim3(1), dim3(256), 0, s, nparts, (const double*)part, (double)A->scale,
                     (double)A->shift * (double)A->n, 1.0 / (64.0 * ngroups), out);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_slq_quadrature(int32_t nb, int32_t m, const double* evals, const double* first_row, const double* scale,
                      int32_t f_kind, double eps, int32_t max_mode, double* out, void* stream) {
  if (nb < 1 || m < 1 || !evals || !first_row || !out || f_kind < 0 || f_kind > 4) return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(slq_kernel, dim3(nb), dim3(64), 0, (hipStream_t)stream, m, evals, first_row, scale, f_kind, eps,
                     max_mode, out);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_basis_combine(int64_t n, int32_t m, int32_t k, const float* V, int64_t ldv, const double* S, int64_t lds_,
                     float* out, int64_t ldo, void* stream) {
  if (n < 1 || m < 1 || k < 1 || !V || !S || !out || ldv < n || ldo < n || lds_ < k) return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(basis_combine_kernel, dim3((unsigned)cdiv(n, 256), (unsigned)cdiv(k, kCombine)), dim3(256), 0,
                     (hipStream_t)stream, n, m, k, V, ldv, S, lds_, out, ldo);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // extern "C"
