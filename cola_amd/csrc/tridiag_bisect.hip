// Symmetric tridiagonal eigensolve by Sturm multisection and inverse iteration, FP64: a
// parallel alternative to the serial QL recurrence for the Lanczos tridiagonal
// (cola/linalg/decompositions/lanczos.py:57 takes xnp.eigh of T).
//
//   * eigenvalues (kernel multisect): L lanes (16 or 64) hunt the j-th smallest eigenvalue
//     together.  Each round lane l evaluates the Sturm count of T - x_l I at
//     x_l = a + (l + 1) h, h = (c - a) / (L + 1) (the number of negative pivots of its LDL^T;
//     a pivot of magnitude below pivmin is replaced by -pivmin), a ballot finds the first
//     point whose count exceeds j, and [a, c] shrinks by L + 1: 9 rounds (L = 64) or 13
//     (L = 16) instead of 52 halvings of plain bisection, each round one O(m) recurrence.
//     The Gershgorin interval starts the search, which stops at a width of
//     2 eps max(|a|, |c|) + eps ||T||, the absolute accuracy the Sturm count has.  The
//     recurrence's division uses the hardware reciprocal refined by two Newton steps: the
//     count is then exact for a relative perturbation of e^2 of a few ulps.
//   * eigenvectors (kernel invit): inverse iteration, x <- (T - lambda I)^-1 x, three solves
//     from a pseudo-random start.  Each solve factorises T - lambda I by Gaussian elimination
//     with partial pivoting while it eliminates the right-hand side (the factorisation is
//     O(m) and reads only LDS, so it is redone per solve instead of kept), storing the
//     pivot reciprocals and the two super-diagonals of U, then back-substitutes; a pivot
//     below eps ||T|| is replaced by +-eps ||T||, which makes the solve grow the
//     eigen-direction.  Eigenvalues closer than 1e-3 ||T|| form a cluster, computed in turn
//     by one thread: a member within 10 eps ||T|| of the previous one is shifted by that
//     much, and after every solve it is orthogonalised against the cluster's earlier vectors
//     (modified Gram-Schmidt).  This is the inverse iteration of Peters & Wilkinson with the
//     clustering rule LAPACK's dstein documents.  One wave per block, whose first tpb lanes
//     take tpb eigenvalues: their factors and vectors (4 m doubles each) live in 64 KB of
//     LDS with the lane index fastest, so tpb = 16 at m = 120 and 2 at m = 1024.
// Outputs use ck_tridiag_eigh_f64's layout: ascending evals, evecs[j * m + c] = component j
// of vector c (its first largest-magnitude component positive), first_row[c] = evecs[c].
#include <float.h>

#include "common.hpp"

using namespace ck;

namespace {

constexpr int kMaxM = 1024;
constexpr int kSolves = 3;
constexpr int kCh = 8;
constexpr int kMsThreads = 256;

struct Bounds {
  double lo, hi, tnorm, pivmin;
};

// d, e, e^2 of problem b into LDS and the Gershgorin interval, ||T|| and pivmin (all threads)
__device__ Bounds load_problem(int m, const float* din, int64_t ldd, const float* ein, int64_t lde, int b,
                               double* d, double* e, double* e2, double* red) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int i = tid; i < m; i += nt) {
    d[i] = (double)din[(int64_t)b * ldd + i];
    e[i] = i < m - 1 ? (double)ein[(int64_t)b * lde + i] : 0.0;
    e2[i] = e[i] * e[i];
  }
  __syncthreads();
  double lo = DBL_MAX, hi = -DBL_MAX, emax = 0.0;
  for (int i = tid; i < m; i += nt) {
    const double rad = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < m - 1 ? fabs(e[i]) : 0.0);
    lo = fmin(lo, d[i] - rad);
    hi = fmax(hi, d[i] + rad);
    emax = fmax(emax, e2[i]);
    // fmin and fmax skip a NaN: a non-finite entry makes the interval infinite instead
    if (!isfinite(d[i]) || !isfinite(e[i])) lo = -INFINITY;
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o));
    hi = fmax(hi, __shfl_xor(hi, o));
    emax = fmax(emax, __shfl_xor(emax, o));
  }
  const int nw = nt >> 6;
  if ((tid & 63) == 0) red[tid >> 6] = lo, red[4 + (tid >> 6)] = hi, red[8 + (tid >> 6)] = emax;
  __syncthreads();
  Bounds B;
  B.lo = red[0], B.hi = red[4];
  emax = red[8];
  for (int w = 1; w < nw; ++w) B.lo = fmin(B.lo, red[w]), B.hi = fmax(B.hi, red[4 + w]), emax = fmax(emax, red[8 + w]);
  B.tnorm = fmax(fabs(B.lo), fabs(B.hi));
  B.pivmin = DBL_MIN * fmax(1.0, emax);
  return B;
}

// 1 / q by the hardware reciprocal and two Newton steps (q is a normal number here)
__device__ inline double recip(double q) {
  double r = __builtin_amdgcn_rcp(q);
  r = fma(r, fma(-q, r, 1.0), r);
  r = fma(r, fma(-q, r, 1.0), r);
  return r;
}

__device__ inline int sturm_count(int m, const double* d, const double* e2, double x, double pivmin) {
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int count = q < 0.0;
  for (int i = 1; i < m; ++i) {
    q = fma(-e2[i - 1], recip(q), d[i] - x);
    if (fabs(q) < pivmin) q = -pivmin;
    count += q < 0.0;
  }
  return count;
}

// L lanes per eigenvalue; grid (ceil(m L / 256), nb)
template <int L>
__global__ __launch_bounds__(kMsThreads) void multisect_kernel(int m, const float* __restrict__ din, int64_t ldd,
                                                               const float* __restrict__ ein, int64_t lde,
                                                               double* __restrict__ evals,
                                                               int32_t* __restrict__ status) {
  __shared__ double d[kMaxM], e[kMaxM], e2[kMaxM], red[12];
  const int b = blockIdx.y, tid = threadIdx.x;
  Bounds B = load_problem(m, din, ldd, ein, lde, b, d, e, e2, red);
  const int j = (blockIdx.x * kMsThreads + tid) / L;
  const int l = tid % L;
  const int shift = (tid & 63) / L * L;  // this group's bits in the wave's ballot
  if (!isfinite(B.tnorm)) {
    if (j < m && l == 0) evals[(int64_t)b * m + j] = NAN;
    if (blockIdx.x == 0 && tid == 0 && status) status[b] = CK_ERR_NAN;
    return;
  }
  const double atol = DBL_EPSILON * fmax(B.tnorm, DBL_MIN);
  // widen the interval so the extreme eigenvalues lie strictly inside it
  const double pad = 2.0 * DBL_EPSILON * B.tnorm * m + 2.0 * B.pivmin;
  double a = B.lo - pad, c = B.hi + pad;
  bool done = j >= m;
  for (int round = 0; round < 64; ++round) {
    if (!done && (c - a <= 2.0 * DBL_EPSILON * fmax(fabs(a), fabs(c)) + atol)) done = true;
    if (!__any(!done)) break;
    const double h = (c - a) / (double)(L + 1);
    const double x = a + (double)(l + 1) * h;
    const bool above = !done && sturm_count(m, d, e2, x, B.pivmin) > j;
    const uint64_t bal = __ballot(above);
    if (!done) {
      const uint64_t mine = (bal >> shift) & (L == 64 ? ~0ull : ((1ull << L) - 1));
      const int f = mine ? __ffsll((unsigned long long)mine) - 1 : L;  // first point with > j eigenvalues below
      const double na = f == 0 ? a : a + (double)f * h;
      const double nc = f == L ? c : a + (double)(f + 1) * h;
      if (!(na > a || nc < c)) done = true;  // no progress: h below the spacing of doubles
      a = na;
      c = nc;
    }
  }
  if (j < m && l == 0) evals[(int64_t)b * m + j] = 0.5 * (a + c);
  if (blockIdx.x == 0 && tid == 0 && status) status[b] = CK_OK;
}

// a start component in [0.5, 1.5), fixed by (i, j), with no structure an eigenvector shares
__device__ inline double start_val(int i, int j) {
  uint32_t h = (uint32_t)(i + 1) * 2654435761u ^ (uint32_t)(j + 1) * 2246822519u;
  h ^= h >> 13;
  h *= 0x5bd1e995u;
  h ^= h >> 15;
  return 0.5 + (double)(h & 0xffffffu) * (1.0 / 16777216.0);
}

// One solve (T - lam I) x_new = s * x_old.  Row r of array a of this thread's scratch is at
// W[(a * m + r) * ts]: array 0 holds the pivot reciprocals of U, arrays 1 and 2 its two
// super-diagonals, array 3 the vector.  Returns ||x_new||^2 and the first largest-magnitude
// component in *big.
__device__ double solve_shifted(int m, const double* d, const double* e, double lam, double s, double tiny,
                                double* W, int64_t ts, double* big) {
  double* R0 = W;
  double* U1 = W + (int64_t)m * ts;
  double* U2 = W + (int64_t)2 * m * ts;
  double* X = W + (int64_t)3 * m * ts;
  // forward elimination; (a, bb) is the active row in columns (i, i+1), r its right-hand side
  double a = d[0] - lam, bb = m > 1 ? e[0] : 0.0;
  double r = X[0] * s;
  for (int i0 = 0; i0 < m - 1; i0 += kCh) {
    double xv[kCh];
#pragma unroll
    for (int u = 0; u < kCh; ++u) xv[u] = X[(int64_t)min(i0 + u + 1, m - 1) * ts];
#pragma unroll
    for (int u = 0; u < kCh; ++u) {
      const int i = i0 + u;
      if (i < m - 1) {
        const double na = e[i], nb = d[i + 1] - lam, nc = i + 1 < m - 1 ? e[i + 1] : 0.0, nr = xv[u] * s;
        const bool swap = fabs(na) > fabs(a);
        double p = swap ? na : a;
        if (fabs(p) < tiny) p = p < 0.0 ? -tiny : tiny;
        const double rp = recip(p);
        R0[(int64_t)i * ts] = rp;
        if (swap) {
          const double mult = a * rp;
          U1[(int64_t)i * ts] = nb;
          U2[(int64_t)i * ts] = nc;
          X[(int64_t)i * ts] = nr;
          a = bb - mult * nb;
          bb = -mult * nc;
          r = r - mult * nr;
        } else {
          const double mult = na * rp;
          U1[(int64_t)i * ts] = bb;
          U2[(int64_t)i * ts] = 0.0;
          X[(int64_t)i * ts] = r;
          a = nb - mult * bb;
          bb = nc;
          r = nr - mult * r;
        }
      }
    }
  }
  {
    double p = a;
    if (fabs(p) < tiny) p = p < 0.0 ? -tiny : tiny;
    R0[(int64_t)(m - 1) * ts] = recip(p);
    U1[(int64_t)(m - 1) * ts] = 0.0;
    U2[(int64_t)(m - 1) * ts] = 0.0;
    X[(int64_t)(m - 1) * ts] = r;
  }
  // back substitution, rows m-1 .. 0
  double x1 = 0.0, x2 = 0.0, nrm2 = 0.0, bg = 0.0;
  for (int i0 = m - 1; i0 >= 0; i0 -= kCh) {
    double p0[kCh], p1[kCh], p2[kCh], y[kCh];
#pragma unroll
    for (int u = 0; u < kCh; ++u) {
      const int64_t i = max(i0 - u, 0);
      p0[u] = R0[i * ts];
      p1[u] = U1[i * ts];
      p2[u] = U2[i * ts];
      y[u] = X[i * ts];
    }
#pragma unroll
    for (int u = 0; u < kCh; ++u) {
      const int i = i0 - u;
      if (i >= 0) {
        const double xi = (y[u] - p1[u] * x1 - p2[u] * x2) * p0[u];
        X[(int64_t)i * ts] = xi;
        x2 = x1;
        x1 = xi;
        nrm2 += xi * xi;
        if (fabs(xi) >= fabs(bg)) bg = xi;  // rows run downwards: ties go to the lowest row
      }
    }
  }
  *big = bg;
  return nrm2;
}

// Inverse iteration for the eigenvalues j0, j0 + 1, ... of problem b that start a cluster at
// j0 (the thread's own index), its factors and vector in W (row r of array a at
// W[(a * m + r) * ts]).  Inlined into each kernel, so that W keeps its address space
// (LDS or global) instead of going through flat accesses.
__device__ __forceinline__ void invit_thread(int m, int b, int j0, const double* d, const double* e, const double* lam,
                             double tnorm, double tiny, double* W, int64_t ts, double* evecs, double* first_row) {
  const double ortol = 1e-3 * tnorm;
  double* Z = evecs + (int64_t)b * m * m;
  double* X = W + (int64_t)3 * m * ts;
  if (j0 >= m || (j0 > 0 && lam[j0] - lam[j0 - 1] <= ortol)) return;  // not a cluster's first
  int j1 = j0 + 1;
  while (j1 < m && lam[j1] - lam[j1 - 1] <= ortol) ++j1;
  double prev = 0.0;
  for (int j = j0; j < j1; ++j) {
    double lj = lam[j];
    if (j > j0) {
      const double pertol = 10.0 * DBL_EPSILON * tnorm;
      if (lj - prev < pertol) lj = prev + pertol;
    }
    prev = lj;
    for (int i = 0; i < m; ++i) X[(int64_t)i * ts] = start_val(i, j);
    double s = 1.0, big = 0.0;
    for (int it = 0; it < kSolves; ++it) {
      double nrm2 = solve_shifted(m, d, e, lj, s, tiny, W, ts, &big);
      if (j > j0) {
        // modified Gram-Schmidt against the cluster's finished vectors (columns j0..j-1): one
        // pass for the first projection, then one per column that subtracts it and forms
        // the next projection (or, after the last column, the norm)
        double dot = 0.0;
        for (int i0 = 0; i0 < m; i0 += kCh) {
          double zv[kCh], xv[kCh];
#pragma unroll
          for (int u = 0; u < kCh; ++u) {
            const int64_t i = min(i0 + u, m - 1);
            zv[u] = Z[i * m + j0];
            xv[u] = X[i * ts];
          }
#pragma unroll
          for (int u = 0; u < kCh; ++u)
            if (i0 + u < m) dot += zv[u] * xv[u];
        }
        for (int k = j0; k < j; ++k) {
          const int kn = min(k + 1, j - 1);
          double nd = 0.0;
          nrm2 = 0.0;
          big = 0.0;
          for (int i0 = 0; i0 < m; i0 += kCh) {
            double zv[kCh], zn[kCh], xv[kCh];
#pragma unroll
            for (int u = 0; u < kCh; ++u) {
              const int64_t i = min(i0 + u, m - 1);
              zv[u] = Z[i * m + k];
              zn[u] = Z[i * m + kn];
              xv[u] = X[i * ts];
            }
#pragma unroll
            for (int u = 0; u < kCh; ++u) {
              if (i0 + u < m) {
                const double xi = xv[u] - dot * zv[u];
                X[(int64_t)(i0 + u) * ts] = xi;
                nd += zn[u] * xi;
                nrm2 += xi * xi;
                if (fabs(xi) > fabs(big)) big = xi;
              }
            }
          }
          dot = nd;
        }
      }
      s = nrm2 > 0.0 ? 1.0 / sqrt(nrm2) : 0.0;
    }
    if (big < 0.0) s = -s;
    for (int i0 = 0; i0 < m; i0 += kCh) {
      double xv[kCh];
#pragma unroll
      for (int u = 0; u < kCh; ++u) xv[u] = X[(int64_t)min(i0 + u, m - 1) * ts];
#pragma unroll
      for (int u = 0; u < kCh; ++u)
        if (i0 + u < m) Z[(int64_t)(i0 + u) * m + j] = xv[u] * s;
    }
    if (first_row) first_row[(int64_t)b * m + j] = X[0] * s;
  }
}

// Two placements of the per-thread scratch (4 m doubles: the pivot reciprocals, two
// super-diagonals and the vector), both with the thread index fastest:
//   invit_lds_kernel: in 64 KB of LDS, 8192 / (4 m) threads per block -- for batches whose
//     blocks fit the CUs at once (one eigensolve of a Lanczos run);
//   invit_global_kernel: in the caller's workspace, 64 threads per block -- for large batches
//     (thousands of small problems), where LDS would hold too few threads per CU.
constexpr int kInvitLds = 8192;

// threads per block with vectors for invit_lds_kernel: the largest power of two <= 64 whose
// scratch fits
inline int invit_tpb(int m) {
  const int fit = kInvitLds / (4 * m);
  int t = 1;
  while (2 * t <= fit && 2 * t <= 64) t *= 2;
  return t;
}
inline int64_t invit_threads(int32_t m) { return ((int64_t)m + 63) / 64 * 64; }

// one wave per block; its first tpb lanes take the eigenvalues blockIdx.x * tpb + lane; grid
// (ceil(m / tpb), nb)
__global__ __launch_bounds__(64) void invit_lds_kernel(int m, int tpb, const float* __restrict__ din, int64_t ldd,
                                                       const float* __restrict__ ein, int64_t lde,
                                                       const double* __restrict__ evals, double* __restrict__ evecs,
                                                       double* __restrict__ first_row) {
  __shared__ double d[kMaxM], e[kMaxM], e2[kMaxM], lam[kMaxM], red[12];
  __shared__ double Ws[kInvitLds];
  const int b = blockIdx.y, tid = threadIdx.x;
  for (int i = tid; i < m; i += 64) lam[i] = evals[(int64_t)b * m + i];
  Bounds B = load_problem(m, din, ldd, ein, lde, b, d, e, e2, red);
  if (!isfinite(B.tnorm)) return;
  const double tnorm = fmax(fabs(lam[0]), fabs(lam[m - 1]));
  const double tiny = DBL_EPSILON * tnorm + B.pivmin;
  if (tid >= tpb) return;
  invit_thread(m, b, blockIdx.x * tpb + tid, d, e, lam, tnorm, tiny, Ws + tid, tpb, evecs, first_row);
}

// 64 threads per block; grid (ceil(m / 64), nb); work: nb x 4 m x invit_threads(m) doubles
__global__ __launch_bounds__(64) void invit_global_kernel(int m, const float* __restrict__ din, int64_t ldd,
                                                          const float* __restrict__ ein, int64_t lde,
                                                          const double* __restrict__ evals,
                                                          double* __restrict__ evecs, double* __restrict__ first_row,
                                                          double* __restrict__ work) {
  __shared__ double d[kMaxM], e[kMaxM], e2[kMaxM], lam[kMaxM], red[12];
  const int b = blockIdx.y, tid = threadIdx.x;
  for (int i = tid; i < m; i += 64) lam[i] = evals[(int64_t)b * m + i];
  Bounds B = load_problem(m, din, ldd, ein, lde, b, d, e, e2, red);
  if (!isfinite(B.tnorm)) return;
  const double tnorm = fmax(fabs(lam[0]), fabs(lam[m - 1]));
  const double tiny = DBL_EPSILON * tnorm + B.pivmin;
  const int64_t ts = (int64_t)gridDim.x * 64;
  const int j0 = blockIdx.x * 64 + tid;
  invit_thread(m, b, j0, d, e, lam, tnorm, tiny, work + (int64_t)b * 4 * m * ts + j0, ts, evecs, first_row);
}

}  // namespace

extern "C" {

size_t ck_tridiag_eigh_bisect_workspace_bytes(int32_t nb, int32_t m) {
  if (nb < 1 || m < 1) return 0;
  return sizeof(double) * (size_t)nb * 4 * (size_t)m * (size_t)invit_threads(m);
}

int ck_tridiag_eigh_bisect_f64(int32_t nb, int32_t m, const float* d, int64_t ldd, const float* e, int64_t lde,
                               double* evals, double* evecs, double* first_row, int32_t* status_out, void* work,
                               size_t work_bytes, void* stream) {
  if (nb < 1 || m < 1 || m > kMaxM || nb > 65535 || !d || !evals || ldd < m || (m > 1 && (!e || lde < m - 1)) ||
      (first_row && !evecs))
    return CK_ERR_BAD_ARG;
  const float* ep = e ? e : d;
  hipStream_t s = (hipStream_t)stream;
  // a wave per eigenvalue while that still leaves few waves in total; 16 lanes otherwise
  if ((int64_t)nb * m <= 4096) {
    const dim3 grid((unsigned)(((int64_t)m * 64 + kMsThreads - 1) / kMsThreads), (unsigned)nb);
    hipLaunchKernelGGL(multisect_kernel<64>, grid, dim3(kMsThreads), 0, s, (int)m, d, ldd, ep, lde, evals, status_out);
  } else {
    const dim3 grid((unsigned)(((int64_t)m * 16 + kMsThreads - 1) / kMsThreads), (unsigned)nb);
    hipLaunchKernelGGL(multisect_kernel<16>, grid, dim3(kMsThreads), 0, s, (int)m, d, ldd, ep, lde, evals, status_out);
  }
  CK_CHECK_LAUNCH();
  if (evecs) {
    // LDS scratch while its one-wave blocks (one per CU) take at most two rounds of the chip
    // (one problem of order 1024: 512 blocks of 2 threads, 47 ms against 61 ms from the
    // workspace); the workspace for larger batches (2000 problems of order 100: 4.7 ms
    // against 12 ms)
    const int tpb = invit_tpb(m);
    const int64_t lds_blocks = (int64_t)nb * ((m + tpb - 1) / tpb);
    const bool global = lds_blocks > 512 && work && work_bytes >= ck_tridiag_eigh_bisect_workspace_bytes(nb, m);
    if (global) {
      const dim3 grid((unsigned)(invit_threads(m) / 64), (unsigned)nb);
      hipLaunchKernelGGL(invit_global_kernel, grid, dim3(64), 0, s, (int)m, d, ldd, ep, lde, evals, evecs, first_row,
                         static_cast<double*>(work));
    } else {
      const dim3 grid((unsigned)((m + tpb - 1) / tpb), (unsigned)nb);
      hipLaunchKernelGGL(invit_lds_kernel, grid, dim3(64), 0, s, (int)m, tpb, d, ldd, ep, lde, evals, evecs,
                         first_row);
    }
    CK_CHECK_LAUNCH();
  }
  return CK_OK;
}

}  // extern "C"
