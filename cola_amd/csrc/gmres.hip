// The small least-squares solve of GMRES on the device.
//
// cola/linalg/inverse/gmres.py:99-121 takes the Arnoldi block (Q, H) of the residual, drops
// H's last row and solves the normal equations (H^T H + D) y = H^T beta e1 of the square
// block, where D has a one on every row of H whose largest |entry| is below 10*tol*max|H|
// ("padded" rows), and then zeroes those rows' y entries.  Those normal equations are the
// least-squares problem  min ||H y - beta e1||^2 + sum_{padded i} y_i^2 , solved here by
// Givens rotations in FP64 without forming H^T H (so the condition number is not squared),
// one workgroup per right-hand side:
//   * only the leading `cols` columns of H can be nonzero (an Arnoldi run that stopped at
//     step k < m leaves the rest zero; their y entries are zero) and only rows < rows
//     (cols + 1, or m when every step ran) meet them;
//   * Givens rotations reduce that upper Hessenberg block to R, with the same rotations
//     applied to g = beta e1;
//   * every padded row i < cols adds the row e_i^T (right-hand side 0), folded into R by
//     rotations over columns i..cols-1;
//   * R y = g by back substitution, then y_i = 0 on padded rows (and on a zero pivot).
#include "common.hpp"

using namespace ck;

namespace {

constexpr int kNT = 256;

constexpr int kMaxRows = 4096;

__device__ inline void givens(double a, double c, double& cs, double& sn) {
  cs = 1.0;
  sn = 0.0;
  if (c != 0.0) {
    const double r = hypot(a, c);
    cs = a / r;
    sn = c / r;
  }
}

__global__ __launch_bounds__(kNT) void hessenberg_lstsq_kernel(int m, const float* __restrict__ H, int64_t ldh,
                                                               int rows, int cols, const float* __restrict__ beta,
                                                               double* __restrict__ work, double* __restrict__ y,
                                                               int64_t ldy, double rtol) {
  __shared__ double s_cs[2];
  __shared__ double red[4];
  __shared__ unsigned char pad[kMaxRows];
  const int b = blockIdx.x, tid = threadIdx.x;
  const float* Hb = H + (int64_t)b * ldh;  // (m+1) x m, row-major
  double* R = work + (int64_t)b * ((int64_t)(rows + 1) * cols + rows + 1);
  double* g = R + (int64_t)rows * cols;    // rows entries
  double* e = g + rows;                    // the extra row folded in for a padded row (cols entries)
  // FP64 copy of the block and of beta e1; the overall largest |entry|
  double amax = 0.0;
  for (int i = tid; i < rows; i += kNT) {
    for (int j = 0; j < cols; ++j) {
      const double v = (double)Hb[(int64_t)i * m + j];
      R[(int64_t)i * cols + j] = v;
      amax = fmax(amax, fabs(v));
    }
    g[i] = i == 0 ? (double)beta[b] : 0.0;
  }
  {
    double v = amax;
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    amax = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    __syncthreads();
  }
  const double thresh = rtol * amax;
  for (int i = tid; i < rows; i += kNT) {
    double rmax = 0.0;
    for (int j = 0; j < cols; ++j) rmax = fmax(rmax, fabs((double)Hb[(int64_t)i * m + j]));
    pad[i] = rmax < thresh ? 1 : 0;
  }
  __syncthreads();
  // Givens QR of the Hessenberg block
  for (int j = 0; j < cols && j + 1 < rows; ++j) {
    if (tid == 0) {
      double cs, sn;
      givens(R[(int64_t)j * cols + j], R[(int64_t)(j + 1) * cols + j], cs, sn);
      s_cs[0] = cs;
      s_cs[1] = sn;
      const double g0 = g[j], g1 = g[j + 1];
      g[j] = cs * g0 + sn * g1;
      g[j + 1] = -sn * g0 + cs * g1;
    }
    __syncthreads();
    const double cs = s_cs[0], sn = s_cs[1];
    for (int l = j + tid; l < cols; l += kNT) {
      const double r0 = R[(int64_t)j * cols + l], r1 = R[(int64_t)(j + 1) * cols + l];
      R[(int64_t)j * cols + l] = cs * r0 + sn * r1;
      R[(int64_t)(j + 1) * cols + l] = l == j ? 0.0 : -sn * r0 + cs * r1;
    }
    __syncthreads();
  }
  // the padded rows' regularisation rows e_i^T; their right-hand side ge (thread 0) starts at 0
  for (int i = 0; i < cols; ++i) {
    if (!pad[i]) continue;
    for (int l = tid; l < cols; l += kNT) e[l] = l == i ? 1.0 : 0.0;
    __syncthreads();
    double ge = 0.0;
    for (int j = i; j < cols; ++j) {
      if (tid == 0) {
        double cs, sn;
        givens(R[(int64_t)j * cols + j], e[j], cs, sn);
        red[0] = cs;
        red[1] = sn;
        const double g0 = g[j];
        g[j] = cs * g0 + sn * ge;
        ge = -sn * g0 + cs * ge;
      }
      __syncthreads();
      const double cs = red[0], sn = red[1];
      for (int l = j + tid; l < cols; l += kNT) {
        const double r0 = R[(int64_t)j * cols + l], r1 = e[l];
        R[(int64_t)j * cols + l] = cs * r0 + sn * r1;
        e[l] = l == j ? 0.0 : -sn * r0 + cs * r1;
      }
      __syncthreads();
    }
  }
  // back substitution
  double* yb = y + (int64_t)b * ldy;
  for (int i = cols - 1; i >= 0; --i) {
    double acc = 0.0;
    for (int l = i + 1 + tid; l < cols; l += kNT) acc += R[(int64_t)i * cols + l] * yb[l];
    acc = wave_sum(acc);
    if ((tid & 63) == 0) red[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) {
      const double s = (red[0] + red[1]) + (red[2] + red[3]);
      const double piv = R[(int64_t)i * cols + i];
      yb[i] = piv != 0.0 ? (g[i] - s) / piv : 0.0;
    }
    __syncthreads();
  }
  for (int i = tid; i < cols; i += kNT)
    if (pad[i]) yb[i] = 0.0;
}

}  // namespace

extern "C" {

size_t ck_hessenberg_lstsq_workspace_bytes(int32_t nb, int32_t rows, int32_t cols) {
  if (nb < 1 || rows < 1 || cols < 0) return 0;
  return sizeof(double) * (size_t)nb * ((size_t)(rows + 1) * cols + rows + 1);
}

int ck_hessenberg_lstsq_f64(int32_t nb, int32_t m, const float* H, int64_t ldh, int32_t rows, int32_t cols,
                            const float* beta, double rtol, void* work, size_t work_bytes, double* y, int64_t ldy,
                            void* stream) {
  if (nb < 1 || m < 1 || !H || !beta || !work || !y || cols < 0 || cols > m || rows < cols || rows > m + 1 ||
      rows < 1 || rows > kMaxRows || ldh < (int64_t)(m + 1) * m || ldy < cols)
    return CK_ERR_BAD_ARG;
  if (work_bytes < ck_hessenberg_lstsq_workspace_bytes(nb, rows, cols)) return CK_ERR_WORKSPACE;
  if (cols == 0) return CK_OK;
  hipLaunchKernelGGL(hessenberg_lstsq_kernel, dim3(nb), dim3(kNT), 0, (hipStream_t)stream, (int)m, H, ldh, (int)rows,
                     (int)cols, beta, static_cast<double*>(work), y, ldy, rtol);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // extern "C"
