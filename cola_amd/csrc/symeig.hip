// Eigen-decomposition of small dense symmetric FP64 matrices on the device.
//
// The Nystrom construction (cola/linalg/preconditioning/preconditioners.py:143-153) ends
// with the SVD of the tall factor B = Y C^-T; its left singular vectors and squared
// singular values are B V diag(sigma^-1) and the eigenpairs (sigma^2, V) of the small Gram
// matrix B^T B, which this kernel computes.  One workgroup per matrix runs the cyclic
// Jacobi method in parallel (round-robin) order:
//   * the matrix sits in LDS (k <= 64, row stride k + 1 so column walks do not conflict);
//   * each round pairs all indices (circle method, one index held fixed), computes every
//     pair's rotation from the current matrix, then applies all rows (J^T A), all columns
//     ((J^T A) J) and the rows of the accumulated E = V^T (E <- J^T E) in global memory;
//   * sweeps (k - 1 rounds, every pair once) repeat until the off-diagonal mass is below
//     (1e-15 ||A||_F)^2 or max_sweeps is reached.
// Output: w descending and E with row j the unit eigenvector of w[j] (E = V^T).
#include "common.hpp"

using namespace ck;

namespace {

constexpr int kNT = 256;
constexpr int kMaxK = 64;
constexpr int kLd = kMaxK + 1;

__device__ inline void pair_of(int round, int i, int m, int& p, int& q) {
  // circle method over m (even) indices: index 0 stays, 1..m-1 rotate by one each round
  const int r = m - 1;
  if (i == 0) {
    p = 0;
    q = 1 + round % r;
  } else {
    p = 1 + (round + i) % r;
    q = 1 + (round - i + r) % r;
  }
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
}

__device__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < kNT / 64; ++i) s += red[i];
  return s;
}

__global__ __launch_bounds__(kNT) void sym_eigh_jacobi_kernel(int k, const double* __restrict__ A, int64_t lda,
                                                              int64_t stride_a, double* __restrict__ w,
                                                              int64_t stride_w, double* __restrict__ E, int64_t lde,
                                                              int64_t stride_e, int max_sweeps,
                                                              int32_t* __restrict__ sweeps_out) {
  __shared__ double a[kMaxK * kLd];
  __shared__ double rc[kMaxK / 2], rs[kMaxK / 2];
  __shared__ double red[kNT / 64];
  __shared__ int perm[kMaxK];
  const int b = blockIdx.x, tid = threadIdx.x;
  const double* Ab = A + (int64_t)b * stride_a;
  double* Eb = E + (int64_t)b * stride_e;
  double* wb = w + (int64_t)b * stride_w;
  const int m = k + (k & 1);  // an odd order gets one idle index
  const int np = m / 2;
  // load the symmetric part, E = I
  double fro = 0.0;
  for (int e = tid; e < k * k; e += kNT) {
    const int i = e / k, j = e % k;
    const double v = 0.5 * (Ab[(int64_t)i * lda + j] + Ab[(int64_t)j * lda + i]);
    a[i * kLd + j] = v;
    fro += v * v;
    Eb[(int64_t)i * lde + j] = i == j ? 1.0 : 0.0;
  }
  fro = block_sum(fro, red);
  const double stop = 1e-30 * fro;
  int sweep = 0;
  bool done = k < 2;
  while (!done && sweep < max_sweeps) {
    for (int round = 0; round < m - 1; ++round) {
      // rotations of this round (Golub & Van Loan, sym.schur2)
      for (int i = tid; i < np; i += kNT) {
        int p, q;
        pair_of(round, i, m, p, q);
        double c = 1.0, s = 0.0;
        if (q < k) {
          const double apq = a[p * kLd + q];
          if (apq != 0.0) {
            const double tau = (a[q * kLd + q] - a[p * kLd + p]) / (2.0 * apq);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
          }
        }
        rc[i] = c;
        rs[i] = s;
      }
      __syncthreads();
      // rows: A <- J^T A, and E <- J^T E
      for (int e = tid; e < np * k; e += kNT) {
        const int i = e / k, j = e % k;
        int p, q;
        pair_of(round, i, m, p, q);
        if (q >= k) continue;
        const double c = rc[i], s = rs[i];
        const double x = a[p * kLd + j], y = a[q * kLd + j];
        a[p * kLd + j] = c * x - s * y;
        a[q * kLd + j] = s * x + c * y;
        const double ex = Eb[(int64_t)p * lde + j], ey = Eb[(int64_t)q * lde + j];
        Eb[(int64_t)p * lde + j] = c * ex - s * ey;
        Eb[(int64_t)q * lde + j] = s * ex + c * ey;
      }
      __syncthreads();
      // columns: A <- A J; the rotated pair's off-diagonal entries are zero by construction
      for (int e = tid; e < np * k; e += kNT) {
        const int i = e / k, r = e % k;
        int p, q;
        pair_of(round, i, m, p, q);
        if (q >= k) continue;
        const double c = rc[i], s = rs[i];
        const double x = a[r * kLd + p], y = a[r * kLd + q];
        a[r * kLd + p] = c * x - s * y;
        a[r * kLd + q] = s * x + c * y;
      }
      __syncthreads();
      for (int i = tid; i < np; i += kNT) {
        int p, q;
        pair_of(round, i, m, p, q);
        if (q < k && rs[i] != 0.0) {
          a[p * kLd + q] = 0.0;
          a[q * kLd + p] = 0.0;
        }
      }
      __syncthreads();
    }
    ++sweep;
    double off = 0.0;
    for (int e = tid; e < k * k; e += kNT) {
      const int i = e / k, j = e % k;
      if (i != j) off += a[i * kLd + j] * a[i * kLd + j];
    }
    off = block_sum(off, red);
    done = off <= stop;
  }
  // descending order of the diagonal; rows of E follow
  if (tid == 0) {
    for (int i = 0; i < k; ++i) perm[i] = i;
    for (int i = 0; i < k; ++i) {
      int best = i;
      for (int j = i + 1; j < k; ++j)
        if (a[perm[j] * kLd + perm[j]] > a[perm[best] * kLd + perm[best]]) best = j;
      const int t = perm[i];
      perm[i] = perm[best];
      perm[best] = t;
    }
  }
  __syncthreads();
  for (int i = tid; i < k; i += kNT) wb[i] = a[perm[i] * kLd + perm[i]];
  __syncthreads();
  // stage E through LDS (the matrix is no longer needed) and write the rows permuted
  for (int e = tid; e < k * k; e += kNT) a[(e / k) * kLd + e % k] = Eb[(int64_t)(e / k) * lde + e % k];
  __syncthreads();
  for (int e = tid; e < k * k; e += kNT) Eb[(int64_t)(e / k) * lde + e % k] = a[perm[e / k] * kLd + e % k];
  if (tid == 0 && sweeps_out) sweeps_out[b] = done ? sweep : -1;
}

}  // namespace

extern "C" {

int ck_sym_eigh_f64(int32_t nb, int32_t k, const double* A, int64_t lda, int64_t stride_a, double* w,
                    int64_t stride_w, double* E, int64_t lde, int64_t stride_e, int32_t max_sweeps, int32_t* sweeps,
                    void* stream) {
  if (nb < 1 || k < 1 || k > kMaxK || !A || !w || !E || lda < k || lde < k || stride_w < k || max_sweeps < 1)
    return CK_ERR_BAD_ARG;
  if (nb > 1 && (stride_a < lda * k || stride_e < lde * k)) return CK_ERR_BAD_ARG;
  hipLaunchKernelGGL(sym_eigh_jacobi_kernel, dim3(nb), dim3(kNT), 0, (hipStream_t)stream, (int)k, A, lda, stride_a,
                     w, stride_w, E, lde, stride_e, (int)max_sweeps, sweeps);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // extern "C"
