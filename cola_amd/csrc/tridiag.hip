// Batched FP64 symmetric tridiagonal eigensolver: implicit QL with Wilkinson
// shifts (Bowdler, Martin, Reinsch & Wilkinson, "The QR and QL algorithms for
// symmetric matrices", Numer. Math. 11 (1968); EISPACK tql2).
//
// One workgroup per problem.  Lane 0 runs the scalar QL recurrence on d/e in
// LDS and logs each sweep's Givens rotations; after a barrier every thread
// applies the sweep to the eigenvector rows it owns (row-parallel, so the
// O(m^3) accumulation is spread over the workgroup).  With first-row-only
// output (stochastic Lanczos quadrature needs tau = Z[0, :]) lane 0 tracks the
// single row itself.
#include <float.h>

#include "common.hpp"

using namespace ck;

namespace {

constexpr int kMaxSweeps = 60;

__global__ __launch_bounds__(256) void tridiag_ql_kernel(int m, const float* __restrict__ din, int64_t ldd,
                                                         const float* __restrict__ ein, int64_t lde,
                                                         double* __restrict__ evals, double* __restrict__ evecs,
                                                         double* __restrict__ first_row,
                                                         int32_t* __restrict__ status) {
  extern __shared__ double sh[];
  double* d = sh;        // m
  double* e = d + m;     // m
  double* cs = e + m;    // m  (rotation cosines of the current sweep)
  double* sn = cs + m;   // m
  double* z0 = sn + m;   // m  (first row of Z when only it is tracked)
  int* perm = reinterpret_cast<int*>(z0 + m);  // m
  __shared__ int ctl_mm, ctl_lo, ctl_fail;
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  double* Z = evecs ? evecs + (int64_t)b * m * m : nullptr;  // row-major Z[j][c]

  for (int i = tid; i < m; i += blockDim.x) {
    d[i] = (double)din[(int64_t)b * ldd + i];
    e[i] = (i < m - 1) ? (double)ein[(int64_t)b * lde + i] : 0.0;
    z0[i] = (i == 0) ? 1.0 : 0.0;
  }
  if (Z)
    for (int64_t t = tid; t < (int64_t)m * m; t += blockDim.x) Z[t] = ((t / m) == (t % m)) ? 1.0 : 0.0;
  if (tid == 0) ctl_fail = 0;
  __syncthreads();

  const int lane = tid & 63;
  for (int l = 0; l < m; ++l) {
    for (int iter = 0;; ++iter) {
      if (tid < 64) {
        // first mm >= l with a negligible off-diagonal (m-1 if none): 64 indices per ballot
        int mm = m - 1;
        for (int base = l; base < m - 1; base += 64) {
          const int idx = base + lane;
          bool small = false;
          if (idx < m - 1) {
            const double dd = fabs(d[idx]) + fabs(d[idx + 1]);
            small = e[idx] == 0.0 || fabs(e[idx]) <= DBL_EPSILON * dd;
          }
          const uint64_t bal = __ballot(small);
          if (bal) {
            mm = base + __ffsll((unsigned long long)bal) - 1;
            break;
          }
        }
        if (tid == 0) {
          ctl_mm = mm;
          ctl_lo = mm;
          if (mm != l) {
            if (iter >= kMaxSweeps) {
              ctl_fail = 1;
              ctl_mm = l;  // give up on this eigenvalue
            } else {
              // Wilkinson shift from the leading 2x2 block
              double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
              double r = hypot(g, 1.0);
              g = d[mm] - d[l] + e[l] / (g + copysign(r, g));
              double s = 1.0, c = 1.0, p = 0.0;
              bool early = false;
              int i = mm - 1;
              // d[i+1] is read by step i and never written before it; d[i] and e[i] are
              // fresh loads, fetched one step ahead so only the arithmetic is serial
              double dip1 = d[mm], di = d[i], ei = e[i];
              double zcarry = Z ? 0.0 : z0[mm];
              for (; i >= l; --i) {
                const double dn = i > l ? d[i - 1] : 0.0;
                const double en = i > l ? e[i - 1] : 0.0;
                const double f = s * ei;
                const double bb = c * ei;
                r = sqrt(fma(f, f, g * g));
                e[i + 1] = r;
                if (r == 0.0) {  // deflation inside the sweep
                  d[i + 1] = dip1 - p;
                  e[mm] = 0.0;
                  early = true;
                  break;
                }
                const double inv = 1.0 / r;
                s = f * inv;
                c = g * inv;
                g = dip1 - p;
                r = (di - g) * s + 2.0 * c * bb;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - bb;
                cs[i] = c;
                sn[i] = s;
                if (!Z) {  // track the first row only (column i+1's value is the carry)
                  const double zi = z0[i];
                  z0[i + 1] = s * zi + c * zcarry;
                  zcarry = c * zi - s * zcarry;
                }
                dip1 = di;
                di = dn;
                ei = en;
              }
              ctl_lo = early ? i + 1 : l;
              if (!Z) z0[early ? i + 1 : l] = zcarry;
              if (!early) {
                d[l] -= p;
                e[l] = g;
                e[mm] = 0.0;
              }
            }
          }
        }
      }
      __syncthreads();
      const int mm = ctl_mm, lo = ctl_lo;
      __syncthreads();  // every lane has read the control words before lane 0 may rewrite them
      if (mm == l) break;
      if (Z) {
        // rotations (i, i+1) for i = mm-1 down to lo, applied to every row.  The value
        // rotation i leaves in column i is the one rotation i-1 reads, so it stays in a
        // register (carry) and each row streams its columns with loads that do not wait
        // on the stores before them (8 loads in flight per row).
        for (int row = tid; row < m; row += blockDim.x) {
          double* zr = Z + (int64_t)row * m;
          double carry = zr[mm];
          int i = mm - 1;
          for (; i - 7 >= lo; i -= 8) {
            double a[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] = zr[i - u];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const double c = cs[i - u], s = sn[i - u];
              zr[i - u + 1] = s * a[u] + c * carry;
              carry = c * a[u] - s * carry;
            }
          }
          for (; i >= lo; --i) {
            const double a = zr[i];
            const double c = cs[i], s = sn[i];
            zr[i + 1] = s * a + c * carry;
            carry = c * a - s * carry;
          }
          zr[lo] = carry;
        }
      }
      __syncthreads();
    }
  }

  // ascending order: each thread ranks its values (ties by index; NaN ranks as +inf so the
  // ranks always form a permutation), perm[rank] = index
  for (int i = tid; i < m; i += blockDim.x) {
    const double di = isnan(d[i]) ? INFINITY : d[i];
    int rank = 0;
    for (int j = 0; j < m; ++j) {
      const double dj = isnan(d[j]) ? INFINITY : d[j];
      rank += (dj < di || (dj == di && j < i)) ? 1 : 0;
    }
    perm[rank] = i;
  }
  if (tid == 0 && status) status[b] = ctl_fail ? CK_ERR_NOCONV : CK_OK;
  __syncthreads();
  for (int c = tid; c < m; c += blockDim.x) {
    evals[(int64_t)b * m + c] = d[perm[c]];
    if (first_row) first_row[(int64_t)b * m + c] = Z ? Z[perm[c]] : z0[perm[c]];
  }
  if (Z) {
    // permute each row in place: new[c] = old[perm[c]], following cycles (same walk on every row)
    __syncthreads();
    for (int row = tid; row < m; row += blockDim.x) {
      double* zr = Z + (int64_t)row * m;
      for (int c0 = 0; c0 < m; ++c0) {
        // c0 leads its cycle iff it is the smallest index on it
        int c = perm[c0];
        bool leader = true;
        while (c != c0) {
          if (c < c0) {
            leader = false;
            break;
          }
          c = perm[c];
        }
        if (!leader) continue;
        const double keep = zr[c0];
        c = c0;
        for (;;) {
          const int nx = perm[c];
          if (nx == c0) {
            zr[c] = keep;
            break;
          }
          zr[c] = zr[nx];
          c = nx;
        }
      }
    }
  }
}

}  // namespace

extern "C" int ck_tridiag_eigh_f64(int32_t nb, int32_t m, const float* d, int64_t ldd, const float* e, int64_t lde,
                                   double* evals, double* evecs, double* first_row, int32_t* status_out,
                                   void* stream) {
  if (nb < 1 || m < 1 || m > 4096 || !d || !evals || ldd < m || (m > 1 && (!e || lde < m - 1))) return CK_ERR_BAD_ARG;
  const size_t lds = (size_t)m * (5 * sizeof(double) + sizeof(int));
  if (lds > 64 * 1024) return CK_ERR_UNSUPPORTED;  // m <= 1489
  hipLaunchKernelGGL(tridiag_ql_kernel, dim3(nb), dim3(256), lds, (hipStream_t)stream, m, d, ldd, e ? e : d, lde,
                     evals, evecs, first_row, status_out);
  CK_CHECK_LAUNCH();
  return CK_OK;
}
