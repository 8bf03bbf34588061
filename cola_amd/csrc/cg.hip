// Batched CG on column blocks with the loop condition evaluated on the device.
//
// Arithmetic and stopping rule follow cola/linalg/inverse/cg.py:94-178 (see the
// ABI header).  Every reduction is two-level with a fixed order: each block of
// (rhs b, chunk) writes one partial, and whoever needs the column total adds
// that column's partials in index order (fp64).  A block of the x/r update
// recomputes its own column's p.Ap from the partials, so no grid-wide barrier
// or atomic is needed; the cross-column decisions (k, done, errors) are made by
// a one-block kernel at the end of each step.
#include "cg_state.hpp"
#include "common.hpp"
#include "ops_internal.hpp"

using namespace ck;

namespace {

constexpr double kSmall = (double)(float)1e-40;  // cg.py:8, as an fp32 constant
constexpr int kChunk = 4096;                      // elements per partial (256 threads x 16)

__host__ __device__ inline int stride_of(int nrhs) { return cg_stride_of(nrhs); }

__host__ __device__ inline CgView view(void* state, int nrhs) { return cg_view(state, nrhs); }

inline int nparts_of(int64_t n) {
  int64_t p = cdiv(n, kChunk);
  if (p < 1) p = 1;
  if (p > 1024) p = 1024;
  return (int)p;
}

// Elements per partial: a multiple of 4 so float4 chunks never straddle blocks.
__host__ __device__ inline int64_t per_of(int64_t n, int P) { return (cdiv(n, P) + 3) / 4 * 4; }

__device__ inline double safe_div(double num, double den) {
  if (fabs(den) < kSmall) den = kSmall;
  return num / den;
}

// Sum of the P partials of column b, fixed order, fp64.
__device__ inline double column_total(const float* __restrict__ part, int P, int b) {
  // each lane adds its partials lane, lane + 64, ... in that order; loads go four at a time
  const float* pb = part + (int64_t)b * P;
  double acc = 0.0;
  int i = threadIdx.x & 63;
  for (; i + 192 < P; i += 256) {
    const float v0 = pb[i], v1 = pb[i + 64], v2 = pb[i + 128], v3 = pb[i + 192];
    acc += (double)v0;
    acc += (double)v1;
    acc += (double)v2;
    acc += (double)v3;
  }
  for (; i < P; i += 64) acc += (double)pb[i];
  return wave_sum(acc);
}

// -------------------------------------------------------------- generic partial sums
// mode 0: sum x*y ; mode 1: sum x*x
template <int MODE>
__global__ __launch_bounds__(256) void partial_kernel(int64_t n, const float* __restrict__ X,
                                                      const float* __restrict__ Yv, int64_t ld,
                                                      float* __restrict__ part, int P, int64_t per, bool v4,
                                                      const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ float red[4];
  const int b = blockIdx.y, blk = blockIdx.x;
  const int64_t i0 = blk * per;
  const int64_t i1 = i0 + per < n ? i0 + per : n;
  const float* x = X + b * ld;
  const float* y = Yv + b * ld;
  float acc = 0.f;
  if (v4) {
    for (int64_t i = i0 + 4 * threadIdx.x; i < i1; i += 1024) {
      const float4 a = *reinterpret_cast<const float4*>(x + i);
      const float4 c = MODE == 0 ? *reinterpret_cast<const float4*>(y + i) : a;
      acc = fmaf(a.x, c.x, acc);
      acc = fmaf(a.y, c.y, acc);
      acc = fmaf(a.z, c.z, acc);
      acc = fmaf(a.w, c.w, acc);
    }
  } else {
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
      const float a = x[i];
      acc = fmaf(a, MODE == 0 ? y[i] : a, acc);
    }
  }
  acc = block_sum<256>(acc, red);
  if (threadIdx.x == 0) part[(int64_t)b * P + blk] = acc;
}

// bn = b / safe(mult)
__global__ __launch_bounds__(256) void scale_cols_kernel(int64_t n, const float* __restrict__ B, int64_t ldb,
                                                         float* __restrict__ Bn, int64_t ldbn,
                                                         const float* __restrict__ mult) {
  const int b = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float m = mult[b];
  if (fabsf(m) < 1e-40f) m = 1e-40f;
  Bn[b * ldbn + i] = B[b * ldb + i] / m;
}

__global__ __launch_bounds__(256) void normalize_final_kernel(int nrhs, const float* __restrict__ part, int P,
                                                              void* state) {
  CgView v = view(state, nrhs);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int b = wave; b < nrhs; b += 4) {
    const double s = column_total(part, P, b);
    if (lane == 0) v.mult[b] = (float)sqrt(s);
  }
}

// p = z and partials of r.r (part1) and r.z (part2, skipped when z == r).
__global__ __launch_bounds__(256) void start_partials_kernel(int64_t n, const float* __restrict__ r,
                                                             const float* __restrict__ z, float* __restrict__ p,
                                                             int64_t ld, float* __restrict__ part1,
                                                             float* __restrict__ part2, int P, int64_t per) {
  __shared__ float red[4];
  const int b = blockIdx.y, blk = blockIdx.x;
  const int64_t i0 = blk * per;
  const int64_t i1 = i0 + per < n ? i0 + per : n;
  float rr = 0.f, rz = 0.f;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
    const float rv = r[b * ld + i], zv = z[b * ld + i];
    p[b * ld + i] = zv;
    rr = fmaf(rv, rv, rr);
    rz = fmaf(rv, zv, rz);
  }
  rr = block_sum<256>(rr, red);
  rz = block_sum<256>(rz, red);
  if (threadIdx.x == 0) {
    part1[(int64_t)b * P + blk] = rr;
    part2[(int64_t)b * P + blk] = rz;
  }
}

__global__ __launch_bounds__(256) void start_final_kernel(int nrhs, const float* __restrict__ part1,
                                                          const float* __restrict__ part2, int P, void* state,
                                                          float tol, int max_iters) {
  __shared__ float rn_lds[1024];
  CgView v = view(state, nrhs);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int b = wave; b < nrhs; b += 4) {
    const double rr = column_total(part1, P, b);
    const double rz = column_total(part2, P, b);
    if (lane == 0) {
      const float rn = (float)sqrt(rr);
      v.rnorm[b] = rn;
      v.gamma[b] = (float)rz;
      v.tol[b] = tol * rn + tol;  // cg.py:101
      v.alpha[b] = 0.f;
      v.beta[b] = 0.f;
      rn_lds[b & 1023] = rn;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float mean = 0.f;
    bool any = false;
    for (int b = 0; b < nrhs; ++b) {
      const float rn = v.rnorm[b];
      mean += rn;
      any = any || (rn > v.tol[b]);
    }
    v.errors[0] = mean / (float)nrhs;
    v.ctrl->k = 0;
    v.ctrl->max_iters = max_iters;
    v.ctrl->nrhs = nrhs;
    v.ctrl->nparts = P;
    v.ctrl->done = (any && 0 < max_iters) ? 0 : 1;  // cg.py:133-138
    // NaN compares false: a non-finite start stops the loop as in the reference
    bool finite = true;
    for (int b = 0; b < nrhs; ++b) finite = finite && isfinite(v.rnorm[b]);
    v.ctrl->status = !finite ? CK_ERR_NAN : (any && v.ctrl->done) ? CK_ERR_MAXITER : CK_OK;
  }
}

// alpha = gamma / p.Ap (safe, frozen -> 0); x += alpha p; r -= alpha Ap; partials of r.r
template <bool V4>
__global__ __launch_bounds__(256) void update_xr_kernel(int64_t n, const float* __restrict__ p,
                                                        const float* __restrict__ Ap, float* __restrict__ x,
                                                        float* __restrict__ r, int64_t ld,
                                                        const float* __restrict__ part0, int np0,
                                                        float* __restrict__ part1, int P, int64_t per, void* state,
                                                        int nrhs) {
  CgView v = view(state, nrhs);
  if (v.ctrl->done) return;
  __shared__ float red[4];
  __shared__ float alpha_s;
  const int b = blockIdx.y, blk = blockIdx.x;
  if (threadIdx.x < 64) {
    const double pap = column_total(part0, np0, b);
    if (threadIdx.x == 0) {
      double a = safe_div((double)v.gamma[b], pap);
      const bool frozen = (double)v.rnorm[b] < kSmall;
      if (frozen) a = 0.0;  // has_converged (cg.py:143-144,161)
      alpha_s = (float)a;
      if (blk == 0) {
        v.alpha[b] = (float)a;
        // the reference divides on regardless; the status records it (same value from every writer)
        if (!frozen && !(pap > 0.0) && v.ctrl->status == CK_OK) v.ctrl->status = CK_ERR_BREAKDOWN;
      }
    }
  }
  __syncthreads();
  const float alpha = alpha_s;
  const int64_t i0 = blk * per;
  const int64_t i1 = i0 + per < n ? i0 + per : n;
  float* xb = x + b * ld;
  float* rb = r + b * ld;
  const float* pb = p + b * ld;
  const float* ab = Ap + b * ld;
  float rr = 0.f;
  if (V4) {
    for (int64_t i = i0 + 4 * threadIdx.x; i < i1; i += 1024) {
      float4 xv = *reinterpret_cast<float4*>(xb + i);
      float4 rv = *reinterpret_cast<float4*>(rb + i);
      const float4 pv = *reinterpret_cast<const float4*>(pb + i);
      const float4 av = *reinterpret_cast<const float4*>(ab + i);
      xv.x = fmaf(alpha, pv.x, xv.x);
      xv.y = fmaf(alpha, pv.y, xv.y);
      xv.z = fmaf(alpha, pv.z, xv.z);
      xv.w = fmaf(alpha, pv.w, xv.w);
      rv.x = fmaf(-alpha, av.x, rv.x);
      rv.y = fmaf(-alpha, av.y, rv.y);
      rv.z = fmaf(-alpha, av.z, rv.z);
      rv.w = fmaf(-alpha, av.w, rv.w);
      *reinterpret_cast<float4*>(xb + i) = xv;
      *reinterpret_cast<float4*>(rb + i) = rv;
      rr = fmaf(rv.x, rv.x, rr);
      rr = fmaf(rv.y, rv.y, rr);
      rr = fmaf(rv.z, rv.z, rr);
      rr = fmaf(rv.w, rv.w, rr);
    }
  } else {
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
      xb[i] = fmaf(alpha, pb[i], xb[i]);
      const float rv = fmaf(-alpha, ab[i], rb[i]);
      rb[i] = rv;
      rr = fmaf(rv, rv, rr);
    }
  }
  rr = block_sum<256>(rr, red);
  if (threadIdx.x == 0) part1[(int64_t)b * P + blk] = rr;
}

// The last step of a refinement correction with the FP64 update folded in:
// x64 += (double)((x + alpha p) * mult), alpha as in update_xr_kernel (x + 0 when the loop
// had already stopped).  The fp32 x and r are left as they were: nothing reads them after.
template <bool V4>
__global__ __launch_bounds__(256) void update_x64_kernel(int64_t n, const float* __restrict__ p,
                                                         const float* __restrict__ x, int64_t ld,
                                                         double* __restrict__ x64, int64_t ldx64,
                                                         const float* __restrict__ part0, int np0, int64_t per,
                                                         void* state, int nrhs) {
  CgView v = view(state, nrhs);
  const bool stopped = v.ctrl->done != 0;
  __shared__ float alpha_s;
  const int b = blockIdx.y, blk = blockIdx.x;
  if (!stopped && threadIdx.x < 64) {
    const double pap = column_total(part0, np0, b);
    if (threadIdx.x == 0) {
      double a = safe_div((double)v.gamma[b], pap);
      if ((double)v.rnorm[b] < kSmall) a = 0.0;
      alpha_s = (float)a;
    }
  }
  __syncthreads();
  const float alpha = stopped ? 0.f : alpha_s;
  const float m = v.mult[b];
  const int64_t i0 = blk * per;
  const int64_t i1 = i0 + per < n ? i0 + per : n;
  const float* xb = x + b * ld;
  const float* pb = p + b * ld;
  double* yb = x64 + b * ldx64;
  if (V4) {
    for (int64_t i = i0 + 4 * threadIdx.x; i < i1; i += 1024) {
      float4 xv = *reinterpret_cast<const float4*>(xb + i);
      if (!stopped) {
        const float4 pv = *reinterpret_cast<const float4*>(pb + i);
        xv.x = fmaf(alpha, pv.x, xv.x);
        xv.y = fmaf(alpha, pv.y, xv.y);
        xv.z = fmaf(alpha, pv.z, xv.z);
        xv.w = fmaf(alpha, pv.w, xv.w);
      }
      yb[i] += (double)(xv.x * m);
      yb[i + 1] += (double)(xv.y * m);
      yb[i + 2] += (double)(xv.z * m);
      yb[i + 3] += (double)(xv.w * m);
    }
  } else {
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
      const float xv = stopped ? xb[i] : fmaf(alpha, pb[i], xb[i]);
      yb[i] += (double)(xv * m);
    }
  }
}

// One PCG step from x = 0 added to an FP64 accumulator: alpha = (r.z) / (z.Az) per column from
// the two partial-sum rows (the generic step's safe division, alpha = 0 for a zero column),
// x64 += (double)(alpha * z).
template <bool V4>
__global__ __launch_bounds__(256) void step1_x64_kernel(int64_t n, const float* __restrict__ z, int64_t ld,
                                                        double* __restrict__ x64, int64_t ldx64,
                                                        const float* __restrict__ prz, const float* __restrict__ pzaz,
                                                        int np, int64_t per) {
  __shared__ float alpha_s;
  const int b = blockIdx.y, blk = blockIdx.x;
  if (threadIdx.x < 64) {
    const double rz = column_total(prz, np, b);
    const double zaz = column_total(pzaz, np, b);
    if (threadIdx.x == 0) alpha_s = (float)safe_div((double)(float)rz, zaz);
  }
  __syncthreads();
  const float alpha = alpha_s;
  const int64_t i0 = blk * per;
  const int64_t i1 = i0 + per < n ? i0 + per : n;
  const float* zb = z + b * ld;
  double* yb = x64 + b * ldx64;
  if (V4) {
    for (int64_t i = i0 + 4 * threadIdx.x; i < i1; i += 1024) {
      const float4 zv = *reinterpret_cast<const float4*>(zb + i);
      yb[i] += (double)(alpha * zv.x);
      yb[i + 1] += (double)(alpha * zv.y);
      yb[i + 2] += (double)(alpha * zv.z);
      yb[i + 3] += (double)(alpha * zv.w);
    }
  } else {
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) yb[i] += (double)(alpha * zb[i]);
  }
}

// End of step: rnorm, gamma1, beta, k, errors, loop condition.
__global__ __launch_bounds__(256) void step_final_kernel(int nrhs, const float* __restrict__ part1,
                                                         const float* __restrict__ part2, int P, int precond,
                                                         void* state) {
  CgView v = view(state, nrhs);
  if (v.ctrl->done) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __shared__ float newrn[1024];
  for (int b = wave; b < nrhs; b += 4) {
    const double rr = column_total(part1, P, b);
    const double rz = precond ? column_total(part2, P, b) : rr;
    if (lane == 0) {
      const bool frozen = (double)v.rnorm[b] < kSmall;  // uses ||r0|| of this step
      double beta = safe_div(rz, (double)v.gamma[b]);
      if (frozen) beta = 0.0;
      v.beta[b] = (float)beta;
      v.gamma[b] = (float)rz;
      newrn[b & 1023] = (float)sqrt(rr);
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nrhs; b += 256) v.rnorm[b] = newrn[b & 1023];
  __syncthreads();
  if (threadIdx.x == 0) {
    const int k = v.ctrl->k + 1;
    float mean = 0.f;
    bool any = false;
    for (int b = 0; b < nrhs; ++b) {
      const float rn = v.rnorm[b];
      mean += rn;
      any = any || (rn > v.tol[b]);
    }
    v.ctrl->k = k;
    v.errors[k] = mean / (float)nrhs;
    if (!(any && k < v.ctrl->max_iters)) {
      v.ctrl->done = 1;
      bool finite = true;
      for (int b = 0; b < nrhs; ++b) finite = finite && isfinite(v.rnorm[b]);
      if (!finite)
        v.ctrl->status = CK_ERR_NAN;
      else if (any && v.ctrl->status == CK_OK)
        v.ctrl->status = CK_ERR_MAXITER;
    }
  }
}

// p = z + beta p
template <bool V4>
__global__ __launch_bounds__(256) void update_p_kernel(int64_t n, const float* __restrict__ z, float* __restrict__ p,
                                                       int64_t ld, void* state, int nrhs) {
  const int b = blockIdx.y;
  const float* zb = z + b * ld;
  float* pb = p + b * ld;
  if (V4) {
    const int64_t i = 4 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    // the vector loads go out before the loop flag and beta are read
    const float4 zv = *reinterpret_cast<const float4*>(zb + i);
    float4 pv = *reinterpret_cast<float4*>(pb + i);
    CgView v = view(state, nrhs);
    if (v.ctrl->done) return;
    const float beta = v.beta[b];
    pv.x = fmaf(beta, pv.x, zv.x);
    pv.y = fmaf(beta, pv.y, zv.y);
    pv.z = fmaf(beta, pv.z, zv.z);
    pv.w = fmaf(beta, pv.w, zv.w);
    *reinterpret_cast<float4*>(pb + i) = pv;
  } else {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    CgView v = view(state, nrhs);
    if (v.ctrl->done) return;
    pb[i] = fmaf(v.beta[b], pb[i], zb[i]);
  }
}

__global__ __launch_bounds__(256) void finish_kernel(int64_t n, float* __restrict__ x, float* __restrict__ r,
                                                     int64_t ld, const float* __restrict__ mult) {
  const int b = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float m = mult[b];
  x[b * ld + i] *= m;
  if (r) r[b * ld + i] *= m;
}

inline bool use_v4(int64_t n, int64_t ld, const void* a, const void* b, const void* c, const void* d) {
  auto al = [](const void* q) { return q == nullptr || (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
  return (n % 4 == 0) && (ld % 4 == 0) && al(a) && al(b) && al(c) && al(d);
}

struct Parts {
  float* p0;
  float* p1;
  float* p2;
  int P;
  int64_t per;
};

inline Parts parts_of(float* partials, int64_t n, int nrhs) {
  Parts q;
  q.P = nparts_of(n);
  q.per = per_of(n, q.P);
  q.p0 = partials;
  q.p1 = partials + (size_t)nrhs * q.P;
  q.p2 = partials + 2 * (size_t)nrhs * q.P;
  return q;
}

int launch_partials(int mode, int64_t n, int nrhs, const float* X, const float* Y, int64_t ld, float* part,
                    const Parts& q, const int* done, hipStream_t s) {
  const bool v4 = use_v4(n, ld, X, Y, nullptr, nullptr);
  dim3 grid(q.P, nrhs);
  if (mode == 0)
    hipLaunchKernelGGL(partial_kernel<0>, grid, dim3(256), 0, s, n, X, Y, ld, part, q.P, q.per, v4, done);
  else
    hipLaunchKernelGGL(partial_kernel<1>, grid, dim3(256), 0, s, n, X, X, ld, part, q.P, q.per, v4, done);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int step_xr(int64_t n, int nrhs, const float* p, const float* Ap, float* x, float* r, int64_t ld, void* state,
            const Parts& q, const float* part0, int np0, hipStream_t s) {
  dim3 grid(q.P, nrhs);
  if (use_v4(n, ld, p, Ap, x, r))
    hipLaunchKernelGGL(update_xr_kernel<true>, grid, dim3(256), 0, s, n, p, Ap, x, r, ld, part0, np0, q.p1, q.P,
                       q.per, state, nrhs);
  else
    hipLaunchKernelGGL(update_xr_kernel<false>, grid, dim3(256), 0, s, n, p, Ap, x, r, ld, part0, np0, q.p1, q.P,
                       q.per, state, nrhs);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

// rz_ready: the preconditioner's last pass already wrote the r.z partials (q.p2)
int step_p(int64_t n, int nrhs, const float* r, const float* z, float* p, int64_t ld, void* state, const Parts& q,
           hipStream_t s, bool rz_ready = false) {
  const int precond = (z != r) ? 1 : 0;
  const int* done = &static_cast<ck_cg_ctrl*>(state)->done;
  if (precond && !rz_ready) {
    int st = launch_partials(0, n, nrhs, r, z, ld, q.p2, q, done, s);
    if (st) return st;
  }
  hipLaunchKernelGGL(step_final_kernel, dim3(1), dim3(256), 0, s, nrhs, q.p1, q.p2, q.P, precond, state);
  if (use_v4(n, ld, z, p, nullptr, nullptr))
    hipLaunchKernelGGL(update_p_kernel<true>, dim3((unsigned)cdiv(n / 4, 256), nrhs), dim3(256), 0, s, n, z, p, ld,
                       state, nrhs);
  else
    hipLaunchKernelGGL(update_p_kernel<false>, dim3((unsigned)cdiv(n, 256), nrhs), dim3(256), 0, s, n, z, p, ld,
                       state, nrhs);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // namespace

extern "C" {

int32_t ck_cg_stride(int32_t nrhs) { return stride_of(nrhs); }

size_t ck_cg_state_bytes(int32_t nrhs, int32_t max_iters) {
  return sizeof(ck_cg_ctrl) + sizeof(float) * ((size_t)CK_CG_ERRORS * stride_of(nrhs) + (size_t)max_iters + 2);
}

int32_t ck_cg_nparts(int64_t n) { return nparts_of(n); }

size_t ck_cg_partials_bytes(int64_t n, int32_t nrhs) { return 3 * sizeof(float) * (size_t)nrhs * nparts_of(n); }

int ck_cg_normalize(int64_t n, int32_t nrhs, const float* b, int64_t ldb, float* bn, int64_t ldbn, void* state,
                    float* partials, void* stream) {
  if (n < 1 || nrhs < 1 || nrhs > 1024 || !b || !bn || !state || !partials || ldb < n || ldbn < n)
    return CK_ERR_BAD_ARG;
  hipStream_t s = (hipStream_t)stream;
  Parts q = parts_of(partials, n, nrhs);
  const bool v4 = use_v4(n, ldb, b, nullptr, nullptr, nullptr);
  hipLaunchKernelGGL(partial_kernel<1>, dim3(q.P, nrhs), dim3(256), 0, s, n, b, b, ldb, q.p1, q.P, q.per, v4,
                     (const int*)nullptr);
  hipLaunchKernelGGL(normalize_final_kernel, dim3(1), dim3(256), 0, s, nrhs, q.p1, q.P, state);
  CgView v = view(state, nrhs);
  hipLaunchKernelGGL(scale_cols_kernel, dim3((unsigned)cdiv(n, 256), nrhs), dim3(256), 0, s, n, b, ldb, bn, ldbn,
                     v.mult);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_cg_start(int64_t n, int32_t nrhs, const float* r, const float* z, float* p, int64_t ld, void* state, float tol,
                int32_t max_iters, float* partials, void* stream) {
  if (n < 1 || nrhs < 1 || nrhs > 1024 || !r || !z || !p || !state || !partials || ld < n || max_iters < 0)
    return CK_ERR_BAD_ARG;
  hipStream_t s = (hipStream_t)stream;
  Parts q = parts_of(partials, n, nrhs);
  hipLaunchKernelGGL(start_partials_kernel, dim3(q.P, nrhs), dim3(256), 0, s, n, r, z, p, ld, q.p1, q.p2, q.P, q.per);
  hipLaunchKernelGGL(start_final_kernel, dim3(1), dim3(256), 0, s, nrhs, q.p1, q.p2, q.P, state, tol, max_iters);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_cg_step_xr(int64_t n, int32_t nrhs, const float* p, const float* Ap, float* x, float* r, int64_t ld,
                  void* state, float* partials, void* stream) {
  if (n < 1 || nrhs < 1 || nrhs > 1024 || !p || !Ap || !x || !r || !state || !partials || ld < n)
    return CK_ERR_BAD_ARG;
  hipStream_t s = (hipStream_t)stream;
  Parts q = parts_of(partials, n, nrhs);
  const int* done = &static_cast<ck_cg_ctrl*>(state)->done;
  int st = launch_partials(0, n, nrhs, p, Ap, ld, q.p0, q, done, s);
  if (st) return st;
  return step_xr(n, nrhs, p, Ap, x, r, ld, state, q, q.p0, q.P, s);
}

int ck_cg_step_p(int64_t n, int32_t nrhs, const float* r, const float* z, float* p, int64_t ld, void* state,
                 float* partials, void* stream) {
  if (n < 1 || nrhs < 1 || nrhs > 1024 || !r || !z || !p || !state || !partials || ld < n) return CK_ERR_BAD_ARG;
  Parts q = parts_of(partials, n, nrhs);
  return step_p(n, nrhs, r, z, p, ld, state, q, (hipStream_t)stream);
}

}  // extern "C"

namespace {

// ------------------------------------------------ PCG step with a Woodbury preconditioner
// For A dense fp32 and P = CK_OP_WOODBURY_INV (rank <= 64, 3..8 rhs) a step runs as
//   dense matvec (p.Ap partials in its epilogue) -> wbf_xr (alpha, x/r update, r.r and
//   F^T r tile partials) -> wbf_core (u = C^-1 F^T r) -> wbf_z (z = scale (r - F u) +
//   shift r, r.z partials) -> step_final -> update_p
// instead of separate partial, F^T r, core, apply and r.z passes.  The Woodbury arithmetic
// after the fp32 loads stays FP64 as in the unfused apply.
struct WbFused {
  float* pap;
  float* rr;
  float* rz;
  double* ftr;
  double* u;
  int T;
  int npap;
};

WbFused wb_fused_carve(void* work, int64_t n, int nrhs, int k) {
  auto al = [](size_t v) { return (v + 255) / 256 * 256; };
  WbFused w;
  w.T = (int)cdiv(n, kWbFusedRows);
  w.npap = dense_xy_nparts(n);
  char* b = static_cast<char*>(work);
  w.pap = reinterpret_cast<float*>(b);
  b += al(4 * (size_t)nrhs * w.npap);
  w.rr = reinterpret_cast<float*>(b);
  b += al(4 * (size_t)nrhs * w.T);
  w.rz = reinterpret_cast<float*>(b);
  b += al(4 * (size_t)nrhs * w.T);
  w.ftr = reinterpret_cast<double*>(b);
  b += al(8 * (size_t)nrhs * k * w.T);
  w.u = reinterpret_cast<double*>(b);
  return w;
}

constexpr int kWbfRows = kWbFusedRows;  // 128: two rows per lane of a wave

// one block per kWbfRows-row tile, all rhs; every global load of the block is issued before
// the first wait (F rows, then p, Ap, x, r), so the tile costs about one memory latency
__global__ __launch_bounds__(256) void wbf_xr_kernel(int64_t n, int nrhs, const float* __restrict__ p,
                                                     const float* __restrict__ Ap, float* __restrict__ x,
                                                     float* __restrict__ r, int64_t ld, const float* __restrict__ pap,
                                                     int npap, float* __restrict__ rr_part,
                                                     const float* __restrict__ F, int64_t lda, int k,
                                                     double* __restrict__ ftr, int T, void* state) {
  constexpr int kRowsPerWave = kWbfRows / 4;
  constexpr int kPer = 8 * kWbfRows / 256;  // (rhs, row) pairs per thread at 8 rhs
  __shared__ float alpha_s[8];
  __shared__ float rs[8][kWbfRows];
  __shared__ double red[4][64][9];
  const int tile = blockIdx.x, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t i0 = (int64_t)tile * kWbfRows;
  // F rows of this wave's quarter of the tile, column = lane.  Every load below reads a valid
  // (clamped) address and its value is used as loaded -- no branch or select waits on it, so
  // all of them are in flight together.  A clamped F value never counts: lanes >= k write no
  // output, and rows >= n meet r = 0 in the product.
  float fv[kRowsPerWave];
  const int cl = lane < k ? lane : k - 1;
#pragma unroll
  for (int j = 0; j < kRowsPerWave; ++j) {
    const int64_t row = i0 + wave * kRowsPerWave + j;
    fv[j] = F[(row < n ? row : n - 1) * lda + cl];
  }
  float pv[kPer], av[kPer], xv[kPer], rv[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int e = threadIdx.x + 256 * j, b = e / kWbfRows, i = e % kWbfRows;
    const int64_t row = i0 + i;
    const int64_t o = (int64_t)(b < nrhs ? b : nrhs - 1) * ld + (row < n ? row : n - 1);
    pv[j] = p[o];
    av[j] = Ap[o];
    xv[j] = x[o];
    rv[j] = r[o];
  }
  // the loop flag is read after the loads are issued (they read valid memory either way)
  CgView v = view(state, nrhs);
  if (v.ctrl->done) return;
  for (int b = wave; b < nrhs; b += 4) {
    const double pap_b = column_total(pap, npap, b);
    if (lane == 0) {
      double a = safe_div((double)v.gamma[b], pap_b);
      const bool frozen = (double)v.rnorm[b] < kSmall;  // has_converged (cg.py:143-144,161)
      if (frozen) a = 0.0;
      alpha_s[b] = (float)a;
      if (tile == 0) {
        v.alpha[b] = (float)a;
        if (!frozen && !(pap_b > 0.0) && v.ctrl->status == CK_OK) v.ctrl->status = CK_ERR_BREAKDOWN;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int e = threadIdx.x + 256 * j, b = e / kWbfRows, i = e % kWbfRows;
    const int64_t row = i0 + i;
    float rn = 0.f;  // rows >= n and rhs >= nrhs stay zero in the product below
    if (b < nrhs && row < n) {
      const float alpha = alpha_s[b];
      const int64_t o = (int64_t)b * ld + row;
      x[o] = fmaf(alpha, pv[j], xv[j]);
      rn = fmaf(-alpha, av[j], rv[j]);
      r[o] = rn;
    }
    rs[b][i] = rn;
  }
  __syncthreads();
  // r.r partials: wave w sums rhs w and w + 4 over the tile
  for (int b = wave; b < nrhs; b += 4) {
    const float a0 = rs[b][lane], a1 = rs[b][lane + 64];
    const float t = wave_sum(fmaf(a0, a0, a1 * a1));
    if (lane == 0) rr_part[(int64_t)b * T + tile] = t;
  }
  // F^T r over the tile: column c = lane, rows split over the four waves, FP64
  double acc[8];
#pragma unroll
  for (int b = 0; b < 8; ++b) acc[b] = 0.0;
#pragma unroll
  for (int j = 0; j < kRowsPerWave; ++j) {
    const double f = (double)fv[j];
    const int i = wave * kRowsPerWave + j;
#pragma unroll
    for (int b = 0; b < 8; ++b) acc[b] = fma(f, (double)rs[b][i], acc[b]);
  }
#pragma unroll
  for (int b = 0; b < 8; ++b) red[wave][lane][b] = acc[b];
  __syncthreads();
  for (int o = threadIdx.x; o < nrhs * 64; o += 256) {
    const int b = o / 64, c = o % 64;
    if (c >= k) continue;
    const double t = (red[0][c][b] + red[1][c][b]) + (red[2][c][b] + red[3][c][b]);
    ftr[((int64_t)b * k + c) * T + tile] = t;
  }
}

// u_b = C^-1 sum_tiles F^T r (tiles in index order within four fixed slices), one block per rhs
__global__ __launch_bounds__(256) void wbf_core_kernel(int k, int nrhs, const double* __restrict__ ftr, int T,
                                                       const double* __restrict__ Cinv, double* __restrict__ u,
                                                       void* state) {
  CgView v = view(state, nrhs);
  if (v.ctrl->done) return;
  __shared__ double red[4][64];
  __shared__ double ts[64];
  const int b = blockIdx.x, c = threadIdx.x & 63, sl = threadIdx.x >> 6;
  if (c < k) {
    const int per = (T + 3) / 4;
    const int t0 = sl * per, t1 = t0 + per < T ? t0 + per : T;
    const double* pc = ftr + ((int64_t)b * k + c) * T;
    double acc = 0.0;
    int t = t0;
    for (; t + 4 <= t1; t += 4) {
      const double a0 = pc[t], a1 = pc[t + 1], a2 = pc[t + 2], a3 = pc[t + 3];
      acc += a0;
      acc += a1;
      acc += a2;
      acc += a3;
    }
    for (; t < t1; ++t) acc += pc[t];
    red[sl][c] = acc;
  }
  __syncthreads();
  if (threadIdx.x < k)
    ts[threadIdx.x] = (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
  __syncthreads();
  if (threadIdx.x < k) {
    const int cc = threadIdx.x;
    double acc = 0.0;
    for (int j = 0; j < k; ++j) acc = fma(Cinv[cc * k + j], ts[j], acc);
    u[(int64_t)b * k + cc] = acc;
  }
}

// z = scale (r - F u) + shift r over the tile (F tile staged in LDS), r.z partials; the F
// tile (float4 loads when rank and row stride allow) and the r values are loaded up front
__global__ __launch_bounds__(256) void wbf_z_kernel(int64_t n, int nrhs, const float* __restrict__ r,
                                                    float* __restrict__ z, int64_t ld, const float* __restrict__ F,
                                                    int64_t lda, int k, const double* __restrict__ u, double scale,
                                                    double shift, float* __restrict__ rz_part, int T, void* state) {
  constexpr int kPer = 8 * kWbfRows / 256;
  __shared__ float Fs[kWbfRows * 65];
  __shared__ double us[8][64];
  __shared__ float prod[8][kWbfRows];
  const int tile = blockIdx.x, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t i0 = (int64_t)tile * kWbfRows;
  float rv[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {  // clamped addresses: no branch around the loads
    const int e = threadIdx.x + 256 * j, b = e / kWbfRows, i = e % kWbfRows;
    const int64_t row = i0 + i;
    rv[j] = r[(int64_t)(b < nrhs ? b : nrhs - 1) * ld + (row < n ? row : n - 1)];
  }
  const bool v4 = (k % 4 == 0) && (lda % 4 == 0) && ((reinterpret_cast<uintptr_t>(F) & 15) == 0);
  if (v4) {
    const int k4 = k / 4;
    constexpr int kMaxF4 = kWbfRows * 16 / 256;  // float4 per thread at rank 64
    float4 f4[kMaxF4];
#pragma unroll
    for (int j = 0; j < kMaxF4; ++j) {
      const int e = threadIdx.x + 256 * j, i = e / k4, c4 = e % k4;
      // clamped address, value used as loaded (rows >= n and entries past the tile are never
      // read back), so the loads stay in flight together
      int64_t row = i0 + (i < kWbfRows ? i : kWbfRows - 1);
      row = row < n ? row : n - 1;
      f4[j] = *reinterpret_cast<const float4*>(F + row * lda + 4 * c4);
    }
#pragma unroll
    for (int j = 0; j < kMaxF4; ++j) {
      const int e = threadIdx.x + 256 * j, i = e / k4, c4 = e % k4;
      if (i < kWbfRows) {
        float* d = Fs + i * 65 + 4 * c4;
        d[0] = f4[j].x;
        d[1] = f4[j].y;
        d[2] = f4[j].z;
        d[3] = f4[j].w;
      }
    }
  } else {
    for (int e = threadIdx.x; e < kWbfRows * k; e += 256) {
      const int i = e / k, c = e % k;
      const int64_t row = i0 + i;
      Fs[i * 65 + c] = row < n ? F[row * lda + c] : 0.f;
    }
  }
  for (int e = threadIdx.x; e < nrhs * k; e += 256) us[e / k][e % k] = u[e];
  CgView v = view(state, nrhs);
  if (v.ctrl->done) return;  // uniform across the block: no thread is left at the barrier
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int e = threadIdx.x + 256 * j, b = e / kWbfRows, i = e % kWbfRows;
    if (b >= nrhs) continue;
    const int64_t row = i0 + i;
    float pr = 0.f;
    if (row < n) {
      double acc = 0.0;
      const float* fr = Fs + i * 65;
      for (int c = 0; c < k; ++c) acc = fma((double)fr[c], us[b][c], acc);
      const double xd = (double)rv[j];
      const float zv = (float)(scale * (xd - acc) + shift * xd);
      z[(int64_t)b * ld + row] = zv;
      pr = rv[j] * zv;
    }
    prod[b][i] = pr;
  }
  __syncthreads();
  for (int b = wave; b < nrhs; b += 4) {
    const float t = wave_sum(prod[b][lane] + prod[b][lane + 64]);
    if (lane == 0) rz_part[(int64_t)b * T + tile] = t;
  }
}

bool wb_fused_ok(const ck_op* A, const ck_op* P, int nrhs, const float* p, int64_t ld, size_t work_bytes) {
  auto al16 = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
  return P && A->kind == CK_OP_DENSE_F32 && P->kind == CK_OP_WOODBURY_INV && nrhs >= 3 && nrhs <= 8 &&
         P->rank >= 1 && P->rank <= 64 && P->lda >= P->rank && P->a && P->aux && A->n % 4 == 0 && A->lda % 4 == 0 &&
         ld % 4 == 0 && al16(A->a) && al16(p) && work_bytes >= wb_fused_bytes(A->n, nrhs, P->rank);
}

// nsteps steps; with last_partial the final one stops after the x/r update
int cg_iterate(const ck_op* A, const ck_op* P, int32_t nrhs, float* x, float* r, float* p, float* Ap, float* z,
               int64_t ld, void* state, float* partials, void* work, size_t work_bytes, int32_t nsteps,
               bool last_partial, double* x64, int64_t ldx64, void* stream) {
  if (!A || nrhs < 1 || nrhs > 1024 || !x || !r || !p || !Ap || !state || !partials || nsteps < 0)
    return CK_ERR_BAD_ARG;
  const int64_t n = A->n;
  if (ld < n) return CK_ERR_BAD_ARG;
  if (P && (P->n != A->n || !z || z == r)) return CK_ERR_BAD_ARG;
  if (!P) z = r;
  hipStream_t s = (hipStream_t)stream;
  Parts q = parts_of(partials, n, nrhs);
  const int* done = &static_cast<ck_cg_ctrl*>(state)->done;
  // the fused epilogue writes nrhs x kron_nparts partials into the p.Ap region (nrhs x P)
  const bool fuse_kron = A->kind == CK_OP_KRON_F32 && A->nfactors == 3 && kron_nparts(A) <= q.P;
  // the 64^3 eigenbasis preconditioner's last pass emits the r.z partials, one per 4096-element
  // slab: the same chunking as the generic partial sums when n = 64^3 (P = 64)
  const bool fuse_eig = P && P->kind == CK_OP_KRON_EIG_F32 && q.P == 64 && q.per == 4096 &&
                        kron64_eligible(P, r, ld, z, ld);
  // dense A with a Woodbury P: the fused step (dense_apply_xy, wbf_*) in P's workspace
  const bool fuse_wb = !fuse_kron && wb_fused_ok(A, P, nrhs, p, ld, work_bytes);
  const bool dense_xy = !fuse_kron && !fuse_wb && A->kind == CK_OP_DENSE_F32 && nrhs >= 3 && nrhs <= 8 && work &&
                        work_bytes >= dense_xy_bytes(n, nrhs) && n % 4 == 0 && A->lda % 4 == 0 && ld % 4 == 0 &&
                        (reinterpret_cast<uintptr_t>(A->a) & 15) == 0 && (reinterpret_cast<uintptr_t>(p) & 15) == 0;
  // CSR A (rows of at most 32 entries on average): p.Ap partials from the SpMV's epilogue, into
  // the head of the workspace like dense_xy
  const bool csr_xy = !fuse_kron && !fuse_wb && !dense_xy && A->kind == CK_OP_CSR_F32 && work &&
                      work_bytes >= csr_xy_bytes(n, nrhs) && (double)A->nnz <= 32.0 * (double)n;
  WbFused W{};
  if (fuse_wb) W = wb_fused_carve(work, n, nrhs, P->rank);
  for (int it = 0; it < nsteps; ++it) {
    int st;
    const bool last = last_partial && it + 1 == nsteps;
    if (fuse_wb) {
      const float* F = static_cast<const float*>(P->a);
      const int k = P->rank;
      st = dense_apply_xy(A, nrhs, p, ld, Ap, ld, done, W.pap, work, work_bytes, s);
      if (st) return st;
      if (last && x64) {
        dim3 grid(q.P, nrhs);
        if (use_v4(n, ld, p, x, nullptr, nullptr) && ldx64 % 4 == 0 && (reinterpret_cast<uintptr_t>(x64) & 31) == 0)
          hipLaunchKernelGGL(update_x64_kernel<true>, grid, dim3(256), 0, s, n, p, x, ld, x64, ldx64,
                             (const float*)W.pap, W.npap, q.per, state, nrhs);
        else
          hipLaunchKernelGGL(update_x64_kernel<false>, grid, dim3(256), 0, s, n, p, x, ld, x64, ldx64,
                             (const float*)W.pap, W.npap, q.per, state, nrhs);
        CK_CHECK_LAUNCH();
        break;
      }
      hipLaunchKernelGGL(wbf_xr_kernel, dim3(W.T), dim3(256), 0, s, n, (int)nrhs, (const float*)p, (const float*)Ap, x,
                         r, ld, (const float*)W.pap, W.npap, W.rr, F, P->lda, k, W.ftr, W.T, state);
      CK_CHECK_LAUNCH();
      if (last) break;
      hipLaunchKernelGGL(wbf_core_kernel, dim3(nrhs), dim3(256), 0, s, k, (int)nrhs, (const double*)W.ftr, W.T,
                         reinterpret_cast<const double*>(P->aux), W.u, state);
      hipLaunchKernelGGL(wbf_z_kernel, dim3(W.T), dim3(256), 0, s, n, (int)nrhs, (const float*)r, z, ld, F, P->lda, k,
                         (const double*)W.u, (double)P->scale, (double)P->shift, W.rz, W.T, state);
      hipLaunchKernelGGL(step_final_kernel, dim3(1), dim3(256), 0, s, (int)nrhs, (const float*)W.rr,
                         (const float*)W.rz, W.T, 1, state);
      if (use_v4(n, ld, z, p, nullptr, nullptr))
        hipLaunchKernelGGL(update_p_kernel<true>, dim3((unsigned)cdiv(n / 4, 256), nrhs), dim3(256), 0, s, n, z, p,
                           ld, state, nrhs);
      else
        hipLaunchKernelGGL(update_p_kernel<false>, dim3((unsigned)cdiv(n, 256), nrhs), dim3(256), 0, s, n, z, p, ld,
                           state, nrhs);
      CK_CHECK_LAUNCH();
      continue;
    }
    int np0 = q.P;
    const float* part0 = q.p0;
    if (fuse_kron) {
      // p.Ap partials come out of the Kronecker epilogue
      np0 = kron_nparts(A);
      st = kron_dispatch(A, nrhs, p, ld, Ap, ld, work, work_bytes, done, q.p0, np0, s);
      if (st) return st;
    } else if (dense_xy) {
      // ... or out of the dense matvec's epilogue, into the head of the workspace (P's apply,
      // which reuses it, runs after the x/r update has read them)
      np0 = dense_xy_nparts(n);
      part0 = static_cast<const float*>(work);
      st = dense_apply_xy(A, nrhs, p, ld, Ap, ld, done, static_cast<float*>(work), work, work_bytes, s);
      if (st) return st;
    } else if (csr_xy) {
      np0 = csr_xy_nparts(n);
      part0 = static_cast<const float*>(work);
      st = csr_apply_xy(A, nrhs, p, ld, Ap, ld, done, static_cast<float*>(work), s);
      if (st) return st;
    } else {
      st = op_apply_internal(A, nrhs, p, ld, Ap, ld, work, work_bytes, done, s);
      if (st) return st;
      st = launch_partials(0, n, nrhs, p, Ap, ld, q.p0, q, done, s);
      if (st) return st;
    }
    if (last && x64) {
      dim3 grid(q.P, nrhs);
      if (use_v4(n, ld, p, x, nullptr, nullptr) && ldx64 % 4 == 0 && (reinterpret_cast<uintptr_t>(x64) & 31) == 0)
        hipLaunchKernelGGL(update_x64_kernel<true>, grid, dim3(256), 0, s, n, p, x, ld, x64, ldx64, part0, np0, q.per,
                           state, nrhs);
      else
        hipLaunchKernelGGL(update_x64_kernel<false>, grid, dim3(256), 0, s, n, p, x, ld, x64, ldx64, part0, np0,
                           q.per, state, nrhs);
      CK_CHECK_LAUNCH();
      break;
    }
    st = step_xr(n, nrhs, p, Ap, x, r, ld, state, q, part0, np0, s);
    if (st) return st;
    if (last) break;
    if (fuse_eig) {
      st = kron64_eig_apply_f32(P, nrhs, r, ld, z, ld, work, work_bytes, done, s, q.p2);
      if (st) return st;
    } else if (P) {
      st = op_apply_internal(P, nrhs, r, ld, z, ld, work, work_bytes, done, s);
      if (st) return st;
    }
    st = step_p(n, nrhs, r, z, p, ld, state, q, s, fuse_eig);
    if (st) return st;
  }
  return CK_OK;
}

}  // namespace

extern "C" {

int ck_cg_iterate(const ck_op* A, const ck_op* P, int32_t nrhs, float* x, float* r, float* p, float* Ap, float* z,
                  int64_t ld, void* state, float* partials, void* work, size_t work_bytes, int32_t nsteps,
                  void* stream) {
  return cg_iterate(A, P, nrhs, x, r, p, Ap, z, ld, state, partials, work, work_bytes, nsteps, false, nullptr, 0,
                    stream);
}

int ck_cg_iterate_final(const ck_op* A, const ck_op* P, int32_t nrhs, float* x, float* r, float* p, float* Ap,
                        float* z, int64_t ld, void* state, float* partials, void* work, size_t work_bytes,
                        int32_t nsteps, void* stream) {
  if (nsteps < 1) return CK_ERR_BAD_ARG;
  return cg_iterate(A, P, nrhs, x, r, p, Ap, z, ld, state, partials, work, work_bytes, nsteps, true, nullptr, 0,
                    stream);
}

int ck_pcg_step1_f64(const ck_op* A, const ck_op* P, int32_t nrhs, const float* r, float* z, float* Az, int64_t ld,
                     double* x64, int64_t ldx64, float* partials, void* work, size_t work_bytes, void* stream) {
  if (!A || !P || nrhs < 1 || nrhs > 1024 || !r || !z || !Az || !x64 || !partials || ld < A->n || ldx64 < A->n)
    return CK_ERR_BAD_ARG;
  const int64_t n = A->n;
  if (P->n != n) return CK_ERR_BAD_ARG;
  Parts q = parts_of(partials, n, nrhs);
  // the fused Kronecker partials (one per 64-column tile of pass B, one per slab of the
  // preconditioner's last pass) line up with the generic 4096-element chunks only at n = 64^3
  const bool ok = A->kind == CK_OP_KRON_F32 && A->nfactors == 3 && kron64_eligible(A, z, ld, Az, ld) &&
                  P->kind == CK_OP_KRON_EIG_F32 && kron64_eligible(P, r, ld, z, ld) && q.P == 64 && q.per == 4096 &&
                  kron_nparts(A) == 64;
  if (!ok) return CK_ERR_UNSUPPORTED;
  hipStream_t s = (hipStream_t)stream;
  int st = kron64_eig_apply_f32(P, nrhs, r, ld, z, ld, work, work_bytes, nullptr, s, q.p2);
  if (st) return st;
  st = kron_dispatch(A, nrhs, z, ld, Az, ld, work, work_bytes, nullptr, q.p0, 64, s);
  if (st) return st;
  dim3 grid(q.P, nrhs);
  if (use_v4(n, ld, z, nullptr, nullptr, nullptr))
    hipLaunchKernelGGL(step1_x64_kernel<true>, grid, dim3(256), 0, s, n, z, ld, x64, ldx64, q.p2, q.p0, q.P, q.per);
  else
    hipLaunchKernelGGL(step1_x64_kernel<false>, grid, dim3(256), 0, s, n, z, ld, x64, ldx64, q.p2, q.p0, q.P, q.per);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

int ck_cg_iterate_final_f64(const ck_op* A, const ck_op* P, int32_t nrhs, float* x, float* r, float* p, float* Ap,
                            float* z, int64_t ld, void* state, float* partials, void* work, size_t work_bytes,
                            int32_t nsteps, double* x64, int64_t ldx64, void* stream) {
  if (nsteps < 1 || !x64 || !A || ldx64 < A->n) return CK_ERR_BAD_ARG;
  return cg_iterate(A, P, nrhs, x, r, p, Ap, z, ld, state, partials, work, work_bytes, nsteps, true, x64, ldx64,
                    stream);
}

int ck_cg_finish(int64_t n, int32_t nrhs, float* x, float* r, int64_t ld, const void* state, void* stream) {
  if (n < 1 || nrhs < 1 || !x || !state || ld < n) return CK_ERR_BAD_ARG;
  CgView v = view(const_cast<void*>(state), nrhs);
  hipLaunchKernelGGL(finish_kernel, dim3((unsigned)cdiv(n, 256), nrhs), dim3(256), 0, (hipStream_t)stream, n, x, r,
                     ld, v.mult);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // extern "C"
