// Persistent fused CG: a whole batched CG solve in one cooperative launch.
//
// Arithmetic and stopping rule are those of cola/linalg/inverse/cg.py:94-178 (column
// normalisation, threshold tol*||r0|| + tol, safe division at 1e-40, frozen columns), the
// loop of cola/utils/torch_tqdm.py:86-90 runs inside the kernel.  Scope: small dense
// operators (y = scale*A x + shift*x) whose solves are launch- and latency-bound, with no
// preconditioner or a diagonal one.
//
// Each workgroup owns a contiguous block of R <= 64 rows.  Its rows of A are staged in LDS
// once (when they fit), and every workgroup keeps the FULL r and p vectors in LDS.  A step:
//   Ap_own = A_own p                   (rows from LDS, p from LDS)
//   publish Ap_own                     (the step's one grid-wide exchange)
//   read all of Ap; p.Ap, alpha        (every workgroup, same order -> same bits)
//   x_own += alpha p_own;  r -= alpha Ap;  gamma' = r.z, ||r||, beta, p = z + beta p
// with the vector updates done redundantly by every workgroup, so a step needs exactly one
// grid-wide exchange and no separate barrier.
//
// The exchange: every Ap entry travels as one 64-bit word (value bits, step tag) written and
// read with agent-scope atomic stores and loads, which the compiler emits as coherent (sc1)
// accesses that do not stop in a per-XCD L2.  Readers load their words and re-read, all
// together, the ones whose tag is still old: the tag is the flag, so a step costs the
// writers' stores landing plus the readers' polling round trips, with no fence, no separate
// flag and no cache maintenance.  The buffers are doubled by step parity: a workgroup can run
// at most one step ahead of the slowest reader (it needs every entry of the current step,
// written after each writer finished reading the previous one), so it never overwrites a word
// still to be read.  Every wait is bounded: after spin_limit polling rounds it gives up,
// records CK_ERR_TIMEOUT in the control block and leaves the loop, so every workgroup ends and
// the launch drains.  All workgroups fit on the device at once (checked by occupancy before
// the launch; a cooperative launch guarantees co-residency, a plain one is the default).
#include <mutex>

#include "cg_state.hpp"
#include "common.hpp"

using namespace ck;

namespace {

constexpr double kSmall = (double)(float)1e-40;  // cg.py:8, as an fp32 constant
constexpr int kNT = 256;
constexpr int kMaxRhs = 4;
constexpr int kMaxBlocks = 256;
constexpr int kMaxRows = 64;              // rows per workgroup
constexpr int kMaxEntries = 7168;         // n * nrhs
constexpr size_t kLdsBudget = 56 * 1024;  // dynamic LDS; the static arrays take < 1 KB more
constexpr int64_t kDefaultSpin = int64_t(1) << 22;

// entries of one column a thread reads back per step (n <= 256 * max_t_of(NB))
__host__ __device__ constexpr int max_t_of(int nb) { return nb == 1 ? 28 : nb == 2 ? 14 : 10; }

struct FusedArgs {
  int n, nn, nrhs, G, R;  // order, padded order (multiple of 4), columns, workgroups, rows per workgroup
  const float* A;
  int64_t lda;
  float scale, shift;
  const float* d;  // diagonal preconditioner (nullptr: none): z = dscale * d * r + dshift * r
  float dscale, dshift;
  const float* b;
  int64_t b_rs, b_cs;
  const float* x0;
  int64_t x0_rs, x0_cs;
  float* x;
  int64_t x_rs, x_cs;
  float* r;
  int64_t r_rs, r_cs;
  void* state;
  float tol;
  int max_iters;
  uint64_t* xbuf;  // 2 x nrhs x nn tagged words
  int64_t spin_limit;
  uint32_t wait_offset;
};

__device__ inline double safe_div(double num, double den) {
  if (fabs(den) < kSmall) den = kSmall;
  return num / den;
}

__device__ inline void put(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint64_t get(const uint64_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void put(int32_t* p, int32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void put(uint32_t* p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline uint32_t get(const uint32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__device__ inline uint64_t pack(float v, uint32_t tag) { return ((uint64_t)tag << 32) | (uint64_t)__float_as_uint(v); }
__device__ inline uint32_t tag_of(uint64_t w) { return (uint32_t)(w >> 32); }
__device__ inline float value_of(uint64_t w) { return __uint_as_float((uint32_t)w); }

// Sum of NV per-thread values over the block, fixed order; every thread gets the totals.
template <int NV>
__device__ inline void block_sums(double (&v)[NV], double* lds /* 4 * NV */) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const double s = wave_sum(v[i]);
    if (lane == 0) lds[wave * NV + i] = s;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = (lds[i] + lds[NV + i]) + (lds[2 * NV + i] + lds[3 * NV + i]);
  __syncthreads();
}

// This thread's entries j = tid + 256 t of every column from the tagged buffer xb (nn words
// per column): the loads of a chunk of sixteen words are issued together, and the words still
// carrying an old tag are re-read together, one round trip per round, until every tag is
// `want` (bounded by spin rounds; after a timeout nothing more is waited for).  Returns false
// on timeout (and sets *fail).
template <int NB, int KT>
__device__ inline bool collect(const uint64_t* xb, int n, int nn, int B, uint32_t want, int64_t spin,
                               float (&out)[NB][KT], int* fail) {
  bool ok = true;
  // entries e = c * KT + t, sixteen at a time: all loads of a chunk are issued together, and
  // the words still carrying an old tag are re-read together, one round trip per round
  constexpr int E = NB * KT, CH = 16;
#pragma unroll
  for (int e0 = 0; e0 < E; e0 += CH) {
    uint64_t raw[CH];
    uint32_t pending = 0;
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int e = e0 + u, c = e / KT, t = e % KT;
      const int j = threadIdx.x + kNT * t;
      if (e < E && c < B && j < n) pending |= 1u << u;
    }
    if (!ok) pending = 0;
    for (int64_t it = 0; pending; ++it) {
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int e = e0 + u, c = e / KT, t = e % KT;
        if (pending & (1u << u)) raw[u] = get(xb + (int64_t)c * nn + threadIdx.x + kNT * t);
      }
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int e = e0 + u, c = e / KT, t = e % KT;
        if ((pending & (1u << u)) && tag_of(raw[u]) == want) {
          out[c][t] = value_of(raw[u]);
          pending &= ~(1u << u);
        }
      }
      if (!pending) break;
      if (it >= spin) {
        ok = false;
        *fail = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  return ok;
}

// NB: columns rounded up to 1, 2 or 4; TILE: own rows of A in LDS
template <bool TILE, int NB>
__global__ __launch_bounds__(kNT) void fused_cg_kernel(FusedArgs a) {
  constexpr int KT = max_t_of(NB);
  extern __shared__ float lds[];
  __shared__ double red[4 * 2 * NB];
  __shared__ float s_alpha[NB], s_beta[NB], s_gamma[NB], s_rnorm[NB], s_tol[NB], s_mult[NB];
  __shared__ float xo[kMaxRows * NB];  // own rows of x, [i * NB + c]
  __shared__ int s_fail;
  const int n = a.n, nn = a.nn, B = a.nrhs, w = blockIdx.x, tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int row0 = w * a.R;
  const int rows = max(0, min(a.R, n - row0));
  float* pv = lds;            // B x nn: full p
  float* rv = pv + B * nn;    // B x nn: full r
  float* At = rv + B * nn;    // R x nn: own rows of A (TILE)
  CgView v = cg_view(a.state, B);
  const bool pre = a.d != nullptr;
  if (tid == 0) s_fail = 0;

  if (TILE) {
    // own rows of A into LDS: eight float4 loads in flight per thread when rows are 16-byte
    // aligned and n is a multiple of 4, else eight scalar loads
    const bool v4 = (n & 3) == 0 && (a.lda & 3) == 0 && (reinterpret_cast<uintptr_t>(a.A) & 15) == 0;
    if (v4) {
      const int q = nn / 4;  // float4 per row
      const int64_t total = (int64_t)a.R * q;
      for (int64_t base = tid; base < total; base += 8 * kNT) {
        float4 buf[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int64_t idx = base + (int64_t)u * kNT;
          const int i = (int)(idx / q), j4 = (int)(idx % q);
          buf[u] = (idx < total && i < rows) ? *reinterpret_cast<const float4*>(a.A + (int64_t)(row0 + i) * a.lda + 4 * j4)
                                              : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int64_t idx = base + (int64_t)u * kNT;
          if (idx < total) *reinterpret_cast<float4*>(At + 4 * idx) = buf[u];
        }
      }
    } else {
      const int64_t total = (int64_t)a.R * nn;
      for (int64_t base = tid; base < total; base += 8 * kNT) {
        float buf[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int64_t idx = base + (int64_t)u * kNT;
          const int i = (int)(idx / nn), j = (int)(idx % nn);
          buf[u] = (idx < total && i < rows && j < n) ? a.A[(int64_t)(row0 + i) * a.lda + j] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int64_t idx = base + (int64_t)u * kNT;
          if (idx < total) At[idx] = buf[u];
        }
      }
    }
  }
  // cg.py:96-97: column norms of b (every workgroup, same order) and r = b / ||b||
  {
    double s[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) s[c] = 0.0;
    for (int j = tid; j < n; j += kNT)
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < B) {
          const double bj = (double)a.b[(int64_t)j * a.b_rs + (int64_t)c * a.b_cs];
          s[c] += bj * bj;
        }
    block_sums<NB>(s, red);
    if (tid < B) s_mult[tid] = (float)sqrt(s[tid]);
    __syncthreads();
  }
  for (int c = 0; c < B; ++c) {
    float m = s_mult[c];
    if (fabsf(m) < 1e-40f) m = 1e-40f;
    for (int j = tid; j < nn; j += kNT) rv[c * nn + j] = j < n ? a.b[(int64_t)j * a.b_rs + (int64_t)c * a.b_cs] / m : 0.f;
  }
  uint32_t sync = 0;
  bool timeout = false;
  // A times the vectors in pv for the own rows, published with tag tg into the buffer xb
  auto gemv_publish = [&](uint64_t* xb, uint32_t tg) {
    for (int i = wave; i < rows; i += 4) {
      const float* arow = TILE ? At + (int64_t)i * nn : a.A + (int64_t)(row0 + i) * a.lda;
      float acc[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) acc[c] = 0.f;
      if (TILE) {
        for (int j = 4 * lane; j < nn; j += 256) {
          const float4 av = *reinterpret_cast<const float4*>(arow + j);
#pragma unroll
          for (int c = 0; c < NB; ++c)
            if (c < B) {
              const float4 p4 = *reinterpret_cast<const float4*>(pv + c * nn + j);
              acc[c] = fmaf(av.x, p4.x, acc[c]);
              acc[c] = fmaf(av.y, p4.y, acc[c]);
              acc[c] = fmaf(av.z, p4.z, acc[c]);
              acc[c] = fmaf(av.w, p4.w, acc[c]);
            }
        }
      } else {
        for (int j = lane; j < n; j += 64) {
          const float aij = arow[j];
#pragma unroll
          for (int c = 0; c < NB; ++c)
            if (c < B) acc[c] = fmaf(aij, pv[c * nn + j], acc[c]);
        }
      }
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < B) {
          const float s = wave_sum(acc[c]);
          if (lane == 0) {
            float y = a.scale * s;
            if (a.shift != 0.f) y = fmaf(a.shift, pv[c * nn + row0 + i], y);
            put(xb + (int64_t)c * nn + row0 + i, pack(y, tg));
          }
        }
    }
  };
  float ap[NB][KT];
  if (a.x0) {
    // cg.py:123: r0 = b_norm - A x0 (x0 as given); one exchange of A x0 through the odd buffer
    // (the first step writes the even one, so a fast workgroup never overwrites it)
    uint64_t* xb = a.xbuf + (int64_t)B * nn;
    for (int c = 0; c < B; ++c)
      for (int j = tid; j < nn; j += kNT) pv[c * nn + j] = j < n ? a.x0[(int64_t)j * a.x0_rs + (int64_t)c * a.x0_cs] : 0.f;
    __syncthreads();
    ++sync;
    gemv_publish(xb, sync);
    collect<NB, KT>(xb, n, nn, B, sync + a.wait_offset, a.spin_limit, ap, &s_fail);
    __syncthreads();
    timeout = s_fail != 0;
    if (!timeout) {
#pragma unroll
      for (int c = 0; c < NB; ++c)
#pragma unroll
        for (int t = 0; t < KT; ++t) {
          const int j = tid + kNT * t;
          if (c < B && j < n) rv[c * nn + j] -= ap[c][t];
        }
      for (int idx = tid; idx < rows * B; idx += kNT) {
        const int i = idx / B, c = idx % B;
        xo[i * NB + c] = pv[c * nn + row0 + i];
      }
    }
  } else {
    for (int idx = tid; idx < rows * B; idx += kNT) xo[(idx / B) * NB + idx % B] = 0.f;
  }
  __syncthreads();
  // cg.py:98-101,122-130: z = P r, p = z, gamma = r.z, threshold tol*||r0|| + tol
  {
    double s[2 * NB];
#pragma unroll
    for (int c = 0; c < 2 * NB; ++c) s[c] = 0.0;
    for (int c = 0; c < B; ++c)
      for (int j = tid; j < nn; j += kNT) {
        const float rj = rv[c * nn + j];
        const float zj = (pre && j < n) ? fmaf(a.dscale * a.d[j], rj, a.dshift * rj) : rj;
        pv[c * nn + j] = zj;
        s[c] += (double)rj * (double)rj;
        s[NB + c] += (double)rj * (double)zj;
      }
    block_sums<2 * NB>(s, red);
    if (tid < B) {
      const float rn = (float)sqrt(s[tid]);
      s_rnorm[tid] = rn;
      s_gamma[tid] = (float)s[NB + tid];
      s_tol[tid] = a.tol * rn + a.tol;
      s_alpha[tid] = 0.f;
      s_beta[tid] = 0.f;
    }
    __syncthreads();
  }
  int k = 0, status = CK_OK;
  bool done;
  {
    float mean = 0.f;
    bool any = false, finite = true;
    for (int c = 0; c < B; ++c) {
      mean += s_rnorm[c];
      any = any || (s_rnorm[c] > s_tol[c]);
      finite = finite && isfinite(s_rnorm[c]);
    }
    if (w == 0 && tid == 0) v.errors[0] = mean / (float)B;
    done = !(any && 0 < a.max_iters);
    if (!finite) status = CK_ERR_NAN;
  }
  while (!done && !timeout) {
    uint64_t* xb = a.xbuf + (int64_t)(k & 1) * B * nn;
    ++sync;
    gemv_publish(xb, sync);
    collect<NB, KT>(xb, n, nn, B, sync + a.wait_offset, a.spin_limit, ap, &s_fail);
    // p.Ap over all n entries (every workgroup, same order) -> alpha (cg.py:143-147)
    double s[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      s[c] = 0.0;
#pragma unroll
      for (int t = 0; t < KT; ++t) {
        const int j = tid + kNT * t;
        if (c < B && j < n) s[c] += (double)pv[c * nn + j] * (double)ap[c][t];
      }
    }
    block_sums<NB>(s, red);
    if (s_fail) {
      timeout = true;
      break;
    }
    if (tid < B) {
      const bool frozen = (double)s_rnorm[tid] < kSmall;
      double al = safe_div((double)s_gamma[tid], s[tid]);
      if (frozen) al = 0.0;
      s_alpha[tid] = (float)al;
    }
    if (tid == 0)
      for (int c = 0; c < B; ++c)
        if (!((double)s_rnorm[c] < kSmall) && !(s[c] > 0.0) && status == CK_OK) status = CK_ERR_BREAKDOWN;
    __syncthreads();
    // x_own += alpha p_own
    for (int idx = tid; idx < rows * B; idx += kNT) {
      const int i = idx / B, c = idx % B;
      xo[i * NB + c] = fmaf(s_alpha[c], pv[c * nn + row0 + i], xo[i * NB + c]);
    }
    // r -= alpha Ap (all entries), gamma' = r.z, ||r||^2
    double q[2 * NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      q[c] = 0.0;
      q[NB + c] = 0.0;
#pragma unroll
      for (int t = 0; t < KT; ++t) {
        const int j = tid + kNT * t;
        if (c < B && j < n) {
          const float rj = fmaf(-s_alpha[c], ap[c][t], rv[c * nn + j]);
          rv[c * nn + j] = rj;
          const float zj = pre ? fmaf(a.dscale * a.d[j], rj, a.dshift * rj) : rj;
          q[c] += (double)rj * (double)rj;
          q[NB + c] += (double)rj * (double)zj;
        }
      }
    }
    block_sums<2 * NB>(q, red);
    // beta = gamma'/gamma (cg.py:150-154), frozen columns keep beta = 0
    if (tid < B) {
      const bool frozen = (double)s_rnorm[tid] < kSmall;
      double be = safe_div(q[NB + tid], (double)s_gamma[tid]);
      if (frozen) be = 0.0;
      s_beta[tid] = (float)be;
      s_gamma[tid] = (float)q[NB + tid];
      s_rnorm[tid] = (float)sqrt(q[tid]);
    }
    __syncthreads();
    for (int c = 0; c < B; ++c) {
      const float be = s_beta[c];
      for (int j = tid; j < n; j += kNT) {
        const float rj = rv[c * nn + j];
        const float zj = pre ? fmaf(a.dscale * a.d[j], rj, a.dshift * rj) : rj;
        pv[c * nn + j] = fmaf(be, pv[c * nn + j], zj);
      }
    }
    // k, errors[k], loop condition (cg.py:133-138)
    ++k;
    float mean = 0.f;
    bool any = false, finite = true;
    for (int c = 0; c < B; ++c) {
      mean += s_rnorm[c];
      any = any || (s_rnorm[c] > s_tol[c]);
      finite = finite && isfinite(s_rnorm[c]);
    }
    if (w == 0 && tid == 0) v.errors[k] = mean / (float)B;
    done = !(any && k < a.max_iters);
    if (done) {
      if (!finite)
        status = CK_ERR_NAN;
      else if (any && status == CK_OK)
        status = CK_ERR_MAXITER;
    }
    __syncthreads();
  }
  if (timeout) {
    if (tid == 0) put(&v.ctrl->status, (int32_t)CK_ERR_TIMEOUT);
    if (w == 0 && tid == 0) {
      v.ctrl->k = k;
      v.ctrl->done = 1;
    }
    return;
  }
  // cg.py:119: x and r back to the scale of b (own rows)
  for (int idx = tid; idx < rows * B; idx += kNT) {
    const int i = idx / B, c = idx % B;
    const float m = s_mult[c];
    a.x[(int64_t)(row0 + i) * a.x_rs + (int64_t)c * a.x_cs] = xo[i * NB + c] * m;
    if (a.r) a.r[(int64_t)(row0 + i) * a.r_rs + (int64_t)c * a.r_cs] = rv[c * nn + row0 + i] * m;
  }
  if (w == 0) {
    if (tid < B) {
      v.mult[tid] = s_mult[tid];
      v.tol[tid] = s_tol[tid];
      v.gamma[tid] = s_gamma[tid];
      v.alpha[tid] = s_alpha[tid];
      v.beta[tid] = s_beta[tid];
      v.rnorm[tid] = s_rnorm[tid];
    }
    if (tid == 0) {
      v.ctrl->k = k;
      v.ctrl->done = 1;
      v.ctrl->max_iters = a.max_iters;
      v.ctrl->nrhs = B;
      v.ctrl->nparts = a.G;
      __hip_atomic_fetch_max(&v.ctrl->status, (int32_t)status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Clears the exchange tags and the status before a solve.
__global__ void fused_reset_kernel(uint64_t* xbuf, int64_t words, void* state) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x)
    xbuf[i] = 0ull;
  if (blockIdx.x == 0 && threadIdx.x == 0) static_cast<ck_cg_ctrl*>(state)->status = CK_OK;
}

struct Plan {
  bool ok, tile;
  int G, R, nn;
  size_t lds;
};

size_t lds_bytes(int nn, int B, int R, bool tile) { return sizeof(float) * ((size_t)2 * B * nn + (tile ? (size_t)R * nn : 0)); }

template <bool TILE>
const void* kernel_nb(int nb) {
  switch (nb) {
    case 1:
      return reinterpret_cast<const void*>(&fused_cg_kernel<TILE, 1>);
    case 2:
      return reinterpret_cast<const void*>(&fused_cg_kernel<TILE, 2>);
    default:
      return reinterpret_cast<const void*>(&fused_cg_kernel<TILE, 4>);
  }
}

const void* kernel_of(bool tile, int nb) { return tile ? kernel_nb<true>(nb) : kernel_nb<false>(nb); }

int nb_of(int B) { return B <= 1 ? 1 : B <= 2 ? 2 : 4; }

// Workgroups that can be resident at once for (kernel, LDS bytes), cached per device.
int max_coresident(bool tile, int nb, size_t lds) {
  struct Entry {
    int dev, key;
    size_t lds;
    int val;
  };
  static Entry cache[64];
  static int used = 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  const int key = (tile ? 16 : 0) + nb;
  for (int i = 0; i < used && i < 64; ++i)
    if (cache[i].dev == dev && cache[i].key == key && cache[i].lds == lds) return cache[i].val;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  int per = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel_of(tile, nb), kNT, lds) != hipSuccess) return 0;
  const int val = per * cus;
  if (used < 64) cache[used++] = Entry{dev, key, lds, val};
  return val;
}

// Workgroups and layout for order n, B columns; nblocks > 0 asks for that many workgroups.
Plan make_plan(int n, int B, int nblocks) {
  Plan p{};
  p.nn = (n + 3) / 4 * 4;
  if (n < 1 || B < 1 || B > kMaxRhs || (int64_t)n * B > kMaxEntries) return p;
  // the per-thread read-back holds max_t_of(nb) entries of each column
  if (n > kNT * max_t_of(nb_of(B))) return p;
  if (lds_bytes(p.nn, B, 0, false) > kLdsBudget) return p;
  // each wave handles rows i, i + 4, ...: at most kMaxRows rows per workgroup
  const int gmin = (int)cdiv(n, kMaxRows);
  int G = nblocks > 0 ? nblocks : (int)cdiv(n, 8);
  if (G > kMaxBlocks) G = kMaxBlocks;
  if (G < gmin) G = gmin;
  if (G > kMaxBlocks || G < 1) return p;
  int R = (int)cdiv(n, G);
  G = (int)cdiv(n, R);
  p.tile = lds_bytes(p.nn, B, R, true) <= kLdsBudget;
  p.G = G;
  p.R = R;
  p.lds = lds_bytes(p.nn, B, R, p.tile);
  p.ok = true;
  return p;
}

size_t work_bytes_of(int nn, int B) { return sizeof(uint64_t) * (size_t)2 * B * nn; }

bool fused_eligible(const ck_op* A, const ck_op* P, int nrhs) {
  if (!A || A->kind != CK_OP_DENSE_F32 || !A->a || A->lda < A->n) return false;
  if (P && !(P->kind == CK_OP_DIAG_F32 && P->a && P->n == A->n)) return false;
  return nrhs >= 1 && nrhs <= kMaxRhs;
}

// The allocation-free entry point's exchange buffer: a file-scope device array, one copy per
// GPU, held until the process ends.  Host code takes its device address once per GPU with
// hipGetSymbolAddress (with that GPU current) and passes that address to the kernels; the
// array's name itself is only a host-side stand-in.  Every call on one GPU shares the array,
// so calls that overlap on one GPU race; the reset kernel clears its tags before each solve.
constexpr int kScratchWords = 8192;  // 64 KiB: 2 * nrhs * (n rounded up to 4) words, i.e. about n * nrhs <= 4096
__device__ __align__(256) uint64_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
uint64_t* g_scratch_ptr[kMaxDevices] = {};

int scratch_of_current_device(uint64_t** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return CK_ERR_HIP;
  if (dev < 0 || dev >= kMaxDevices) return CK_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CK_ERR_HIP;
    g_scratch_ptr[dev] = static_cast<uint64_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return CK_OK;
}

// The solve with its exchange buffer at xbuf (xbuf_bytes available); shared by the entry points.
int fused_launch(const ck_op* A, const ck_op* P, int32_t nrhs, const float* b, int64_t b_rs, int64_t b_cs,
                 const float* x0, int64_t x0_rs, int64_t x0_cs, float* x, int64_t x_rs, int64_t x_cs, float* r,
                 int64_t r_rs, int64_t r_cs, void* state, float tol, int32_t max_iters, uint64_t* xbuf,
                 size_t xbuf_bytes, int32_t nblocks, int64_t spin_limit, int32_t wait_offset, int32_t cooperative,
                 void* stream) {
  if (!A || !b || !x || !state || !xbuf || max_iters < 0 || nblocks < 0 || wait_offset < 0) return CK_ERR_BAD_ARG;
  if (!fused_eligible(A, P, nrhs)) return CK_ERR_UNSUPPORTED;
  const int n = A->n;
  Plan pl = make_plan(n, nrhs, nblocks);
  if (!pl.ok) return CK_ERR_UNSUPPORTED;
  if (xbuf_bytes < work_bytes_of(pl.nn, nrhs)) return CK_ERR_WORKSPACE;
  const int nb = nb_of(nrhs);
  if (pl.G > max_coresident(pl.tile, nb, pl.lds)) return CK_ERR_UNSUPPORTED;
  FusedArgs a{};
  a.n = n;
  a.nn = pl.nn;
  a.nrhs = nrhs;
  a.G = pl.G;
  a.R = pl.R;
  a.A = static_cast<const float*>(A->a);
  a.lda = A->lda;
  a.scale = A->scale;
  a.shift = A->shift;
  if (P) {
    a.d = static_cast<const float*>(P->a);
    a.dscale = P->scale;
    a.dshift = P->shift;
  }
  a.b = b;
  a.b_rs = b_rs;
  a.b_cs = b_cs;
  a.x0 = x0;
  a.x0_rs = x0_rs;
  a.x0_cs = x0_cs;
  a.x = x;
  a.x_rs = x_rs;
  a.x_cs = x_cs;
  a.r = r;
  a.r_rs = r_rs;
  a.r_cs = r_cs;
  a.state = state;
  a.tol = tol;
  a.max_iters = max_iters;
  a.xbuf = xbuf;
  a.spin_limit = spin_limit > 0 ? spin_limit : kDefaultSpin;
  a.wait_offset = (uint32_t)wait_offset;
  hipStream_t s = (hipStream_t)stream;
  const int64_t words = (int64_t)2 * nrhs * pl.nn;
  const unsigned rblocks = (unsigned)(cdiv(words, 1024) < 64 ? cdiv(words, 1024) : 64);
  hipLaunchKernelGGL(fused_reset_kernel, dim3(rblocks), dim3(256), 0, s, a.xbuf, words, state);
  CK_CHECK_LAUNCH();
  void* args[] = {&a};
  // every workgroup fits at once (checked above); a cooperative launch also guarantees that
  // they are resident together, a plain one costs about 20 us less per solve -- with the
  // bounded waits, a workgroup that started late makes the others wait, never hang
  const hipError_t e = cooperative
                           ? hipLaunchCooperativeKernel(kernel_of(pl.tile, nb), dim3(pl.G), dim3(kNT), args,
                                                        (unsigned)pl.lds, s)
                           : hipLaunchKernel(kernel_of(pl.tile, nb), dim3(pl.G), dim3(kNT), args, (unsigned)pl.lds, s);
  if (e != hipSuccess) return CK_ERR_HIP;
  return CK_OK;
}

}  // namespace

extern "C" {

size_t ck_cg_fused_workspace_bytes(int64_t n, int32_t nrhs) {
  if (n < 1 || nrhs < 1) return 0;
  return work_bytes_of((int)((n + 3) / 4 * 4), nrhs);
}

int ck_cg_fused_supported(const ck_op* A, const ck_op* P, int32_t nrhs) {
  if (!fused_eligible(A, P, nrhs)) return CK_ERR_UNSUPPORTED;
  Plan p = make_plan(A->n, nrhs, 0);
  return p.ok ? CK_OK : CK_ERR_UNSUPPORTED;
}

int ck_cg_fused(const ck_op* A, const ck_op* P, int32_t nrhs, const float* b, int64_t b_rs, int64_t b_cs,
                const float* x0, int64_t x0_rs, int64_t x0_cs, float* x, int64_t x_rs, int64_t x_cs, float* r,
                int64_t r_rs, int64_t r_cs, void* state, float tol, int32_t max_iters, void* work,
                size_t work_bytes, int32_t nblocks, int64_t spin_limit, int32_t wait_offset, int32_t cooperative,
                void* stream) {
  if (!work || (reinterpret_cast<uintptr_t>(work) & 7) != 0) return !work ? CK_ERR_BAD_ARG : CK_ERR_WORKSPACE;
  return fused_launch(A, P, nrhs, b, b_rs, b_cs, x0, x0_rs, x0_cs, x, x_rs, x_cs, r, r_rs, r_cs, state, tol, max_iters,
                      static_cast<uint64_t*>(work), work_bytes, nblocks, spin_limit, wait_offset, cooperative, stream);
}

int ck_cg_fused_noalloc(const ck_op* A, const ck_op* P, int32_t nrhs, const float* b, int64_t b_rs, int64_t b_cs,
                        const float* x0, int64_t x0_rs, int64_t x0_cs, float* x, int64_t x_rs, int64_t x_cs, float* r,
                        int64_t r_rs, int64_t r_cs, void* state, float tol, int32_t max_iters, int32_t nblocks,
                        int64_t spin_limit, int32_t wait_offset, int32_t cooperative, void* stream) {
  uint64_t* ptr = nullptr;
  const int st = scratch_of_current_device(&ptr);
  if (st != CK_OK) return st;
  return fused_launch(A, P, nrhs, b, b_rs, b_cs, x0, x0_rs, x0_cs, x, x_rs, x_cs, r, r_rs, r_cs, state, tol, max_iters,
                      ptr, sizeof(scratch), nblocks, spin_limit, wait_offset, cooperative, stream);
}

size_t ck_cg_fused_noalloc_max_bytes(void) { return sizeof(scratch); }

}  // extern "C"
