// Batched FP64 symmetric tridiagonal eigensolver: implicit QL with Wilkinson
// shifts (Bowdler, Martin, Reinsch & Wilkinson, "The QR and QL algorithms for
// symmetric matrices", Numer. Math. 11 (1968); EISPACK tql2).
//
// One workgroup per problem.  Lane 0 runs the scalar QL recurrence on d/e in
// LDS and logs each sweep's Givens rotations; the sweep is applied to the
// eigenvector rows row-parallel.  With first-row-only output (stochastic
// Lanczos quadrature needs tau = Z[0, :]) lane 0 tracks the single row itself
// (tridiag_ql_kernel); with eigenvectors, sweep k+1's recurrence overlaps the
// application of sweep k (tridiag_ql_vec_kernel).
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

// Eigenvector variant with the two halves of a QL step overlapped.  Wave 0 produces sweeps:
// it finds the next unreduced block with a wave-wide ballot and lane 0 runs the scalar
// recurrence, logging the sweep's rotations into one of two buffers.  Waves 1-3 meanwhile
// apply the previous sweep's rotations to the rows of Z (row-parallel).  One barrier per
// step; the control words are double-buffered with the rotations, so every thread reads the
// loop condition from a buffer nobody writes until two steps later.  With ZLDS the m x m
// eigenvector matrix lives in LDS and is written out (permuted) at the end.  The arithmetic
// of every rotation and the order in which each row receives them are those of
// tridiag_ql_kernel.
template <bool ZLDS>
__global__ __launch_bounds__(256) void tridiag_ql_vec_kernel(int m, const float* __restrict__ din, int64_t ldd,
                                                             const float* __restrict__ ein, int64_t lde,
                                                             double* __restrict__ evals, double* __restrict__ evecs,
                                                             double* __restrict__ first_row,
                                                             int32_t* __restrict__ status) {
  extern __shared__ double sh[];
  double* d = sh;  // m
  double* e = d + m;
  double* csb[2] = {e + m, e + 2 * m};
  double* snb[2] = {e + 3 * m, e + 4 * m};
  int* perm = reinterpret_cast<int*>(e + 5 * m);  // m ints, padded to a multiple of 2
  double* Zl = reinterpret_cast<double*>(perm + ((m + 1) & ~1));
  __shared__ int s_mm[2], s_lo[2], s_valid[2], s_fin[2], s_fail;
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  double* Z = ZLDS ? Zl : evecs + (int64_t)b * m * m;  // row-major Z[j][c]

  for (int i = tid; i < m; i += blockDim.x) {
    d[i] = (double)din[(int64_t)b * ldd + i];
    e[i] = (i < m - 1) ? (double)ein[(int64_t)b * lde + i] : 0.0;
  }
  for (int64_t t = tid; t < (int64_t)m * m; t += blockDim.x) Z[t] = ((t / m) == (t % m)) ? 1.0 : 0.0;
  if (tid == 0) {
    s_fail = 0;
    s_valid[0] = s_valid[1] = 0;
    s_fin[0] = s_fin[1] = 0;
  }
  __syncthreads();

  int l = 0, iter = 0;  // producer state (wave 0; every lane of it holds the same values)
  int buf = 0;          // buffer written by this step's producer; consumers apply buf ^ 1
  for (;;) {
    if (tid < 64) {
      bool produced = false, fin = false;
      for (;;) {
        if (l >= m) {
          fin = true;
          break;
        }
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
        if (mm == l) {
          ++l;
          iter = 0;
          continue;
        }
        if (iter >= kMaxSweeps) {  // give up on this eigenvalue
          if (lane == 0) s_fail = 1;
          ++l;
          iter = 0;
          continue;
        }
        if (lane == 0) {
          double* cs = csb[buf];
          double* sn = snb[buf];
          double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
          double r = hypot(g, 1.0);
          g = d[mm] - d[l] + e[l] / (g + copysign(r, g));
          double s = 1.0, c = 1.0, p = 0.0;
          bool early = false;
          int i = mm - 1;
          double dip1 = d[mm], di = d[i], ei = e[i];
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
            dip1 = di;
            di = dn;
            ei = en;
          }
          if (!early) {
            d[l] -= p;
            e[l] = g;
            e[mm] = 0.0;
          }
          s_mm[buf] = mm;
          s_lo[buf] = early ? i + 1 : l;
        }
        ++iter;
        produced = true;
        break;
      }
      if (lane == 0) {
        s_valid[buf] = produced ? 1 : 0;
        s_fin[buf] = fin ? 1 : 0;
      }
    } else if (s_valid[buf ^ 1]) {
      // rotations (i, i+1) for i = mm-1 down to lo of the previous sweep, applied to every row
      const int mm = s_mm[buf ^ 1], lo = s_lo[buf ^ 1];
      const double* cs = csb[buf ^ 1];
      const double* sn = snb[buf ^ 1];
      for (int row = tid - 64; row < m; row += blockDim.x - 64) {
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
    // nothing produced and the previous sweep applied in this step: done (buffer `buf` is not
    // rewritten before the barrier of the step after next, so all threads read the same words)
    if (s_fin[buf] && !s_valid[buf]) break;
    buf ^= 1;
  }

  for (int i = tid; i < m; i += blockDim.x) {
    const double di = isnan(d[i]) ? INFINITY : d[i];
    int rank = 0;
    for (int j = 0; j < m; ++j) {
      const double dj = isnan(d[j]) ? INFINITY : d[j];
      rank += (dj < di || (dj == di && j < i)) ? 1 : 0;
    }
    perm[rank] = i;
  }
  if (tid == 0 && status) status[b] = s_fail ? CK_ERR_NOCONV : CK_OK;
  __syncthreads();
  for (int c = tid; c < m; c += blockDim.x) {
    evals[(int64_t)b * m + c] = d[perm[c]];
    if (first_row) first_row[(int64_t)b * m + c] = Z[perm[c]];
  }
  if (ZLDS) {
    double* out = evecs + (int64_t)b * m * m;
    for (int64_t t = tid; t < (int64_t)m * m; t += blockDim.x) {
      const int row = (int)(t / m), c = (int)(t % m);
      out[t] = Zl[(int64_t)row * m + perm[c]];
    }
  } else {
    // permute each row in place: new[c] = old[perm[c]], following cycles (same walk on every row)
    __syncthreads();
    for (int row = tid; row < m; row += blockDim.x) {
      double* zr = Z + (int64_t)row * m;
      for (int c0 = 0; c0 < m; ++c0) {
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

size_t vec_lds_bytes(int m, bool zlds) {
  return (size_t)m * 6 * sizeof(double) + (size_t)((m + 1) & ~1) * sizeof(int) +
         (zlds ? (size_t)m * m * sizeof(double) : 0);
}

}  // namespace

extern "C" int ck_tridiag_eigh_f64(int32_t nb, int32_t m, const float* d, int64_t ldd, const float* e, int64_t lde,
                                   double* evals, double* evecs, double* first_row, int32_t* status_out,
                                   void* stream) {
  if (nb < 1 || m < 1 || m > 4096 || !d || !evals || ldd < m || (m > 1 && (!e || lde < m - 1))) return CK_ERR_BAD_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (evecs) {
    // eigenvectors: overlapped sweeps; Z in LDS when it fits the device's per-workgroup limit
    static int zlds_limit = -1;  // bytes a workgroup may use (opt-in), queried once
    if (zlds_limit < 0) {
      // per-workgroup LDS: the device maximum (160 KB on gfx950), or the opt-in figure if larger
      int dev = 0, maxb = 0, optin = 0;
      if (hipGetDevice(&dev) == hipSuccess) {
        if (hipDeviceGetAttribute(&maxb, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) maxb = 0;
        if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess) optin = 0;
      }
      if (optin < maxb) optin = maxb;
      if (optin < 64 * 1024) optin = 64 * 1024;
      // the kernel's own static LDS (control words) comes out of the same budget
      zlds_limit = optin - 1024;
      if (zlds_limit > 64 * 1024 &&
          hipFuncSetAttribute(reinterpret_cast<const void*>(&tridiag_ql_vec_kernel<true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, zlds_limit) != hipSuccess)
        zlds_limit = 64 * 1024 - 1024;
      (void)hipGetLastError();
    }
    const size_t lz = vec_lds_bytes(m, true);
    if (lz <= (size_t)zlds_limit) {
      hipLaunchKernelGGL(tridiag_ql_vec_kernel<true>, dim3(nb), dim3(256), lz, s, m, d, ldd, e ? e : d, lde, evals,
                         evecs, first_row, status_out);
    } else {
      const size_t lg = vec_lds_bytes(m, false);
      if (lg > 64 * 1024) return CK_ERR_UNSUPPORTED;
      hipLaunchKernelGGL(tridiag_ql_vec_kernel<false>, dim3(nb), dim3(256), lg, s, m, d, ldd, e ? e : d, lde, evals,
                         evecs, first_row, status_out);
    }
    CK_CHECK_LAUNCH();
    return CK_OK;
  }
  const size_t lds = (size_t)m * (5 * sizeof(double) + sizeof(int));
  if (lds > 64 * 1024) return CK_ERR_UNSUPPORTED;  // m <= 1489
  hipLaunchKernelGGL(tridiag_ql_kernel, dim3(nb), dim3(256), lds, s, m, d, ldd, e ? e : d, lde, evals, evecs,
                     first_row, status_out);
  CK_CHECK_LAUNCH();
  return CK_OK;
}
