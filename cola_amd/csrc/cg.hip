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
  CgView v = view(state, nrhs);
  if (v.ctrl->done) return;
  const int b = blockIdx.y;
  const float beta = v.beta[b];
  const float* zb = z + b * ld;
  float* pb = p + b * ld;
  if (V4) {
    const int64_t i = 4 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const float4 zv = *reinterpret_cast<const float4*>(zb + i);
    float4 pv = *reinterpret_cast<float4*>(pb + i);
    pv.x = fmaf(beta, pv.x, zv.x);
    pv.y = fmaf(beta, pv.y, zv.y);
    pv.z = fmaf(beta, pv.z, zv.z);
    pv.w = fmaf(beta, pv.w, zv.w);
    *reinterpret_cast<float4*>(pb + i) = pv;
  } else {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pb[i] = fmaf(beta, pb[i], zb[i]);
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
  for (int it = 0; it < nsteps; ++it) {
    int st;
    const bool last = last_partial && it + 1 == nsteps;
    int np0 = q.P;
    if (fuse_kron) {
      // p.Ap partials come out of the Kronecker epilogue
      np0 = kron_nparts(A);
      st = kron_dispatch(A, nrhs, p, ld, Ap, ld, work, work_bytes, done, q.p0, np0, s);
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
        hipLaunchKernelGGL(update_x64_kernel<true>, grid, dim3(256), 0, s, n, p, x, ld, x64, ldx64, q.p0, np0, q.per,
                           state, nrhs);
      else
        hipLaunchKernelGGL(update_x64_kernel<false>, grid, dim3(256), 0, s, n, p, x, ld, x64, ldx64, q.p0, np0, q.per,
                           state, nrhs);
      CK_CHECK_LAUNCH();
      break;
    }
    st = step_xr(n, nrhs, p, Ap, x, r, ld, state, q, q.p0, np0, s);
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
