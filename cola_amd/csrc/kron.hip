// Kronecker matvecs with 64x64 factors on MFMA (fp32: v_mfma_f32_16x16x4f32,
// fp64: v_mfma_f64_16x16x4f64).
//
// Layout: vector b of a column block is a 64x64x64 tensor X[i1][i2][i3].
//   pass A (one block per (b, i1) slab S = X[b][i1] as a 64x64 matrix):
//       U = K2 S K3^T          (modes 2 and 3 fused; the 64x64 intermediate
//                               stays in LDS, transposed)
//   pass B (one block per (b, 64-column tile of the i2*i3 plane)):
//       Y = scale * K1 U + shift * X   (optionally times an n-vector d, and
//                                        emitting per-block partials of X.Y)
//
// MFMA operand trick: the sum over k is order-free, so lane l of a 16x16x4
// step takes k = 16*(l/16) + ks (ks = 0..15) instead of k = 4*ks + l/16.  Each
// lane's A and B operands for the whole 64-long reduction are then 16
// contiguous values: four 16-byte loads straight from global memory (factors
// and slab rows) or from LDS, and the 16 MFMAs of a tile issue back to back.
#include "common.hpp"
#include "ops_internal.hpp"

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

namespace {

template <typename T>
struct MF;
// Accumulator layout (measured): lane l, item r of a 16x16 result holds
// D[row(l/16, r)][l % 16] with row = 4*(l/16) + r for f32 and (l/16) + 4*r for f64.
template <>
struct MF<float> {
  using acc = floatx4;
  __device__ __forceinline__ static int row(int g, int r) { return 4 * g + r; }
  __device__ __forceinline__ static acc step(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template <>
struct MF<double> {
  using acc = doublex4;
  __device__ __forceinline__ static int row(int g, int r) { return g + 4 * r; }
  __device__ __forceinline__ static acc step(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

// 16 contiguous elements (16-byte aligned) into registers
__device__ __forceinline__ void load16(const float* __restrict__ p, float* r) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float4 v = *reinterpret_cast<const float4*>(p + 4 * q);
    r[4 * q] = v.x;
    r[4 * q + 1] = v.y;
    r[4 * q + 2] = v.z;
    r[4 * q + 3] = v.w;
  }
}
__device__ __forceinline__ void load16(const double* __restrict__ p, double* r) {
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const double2 v = *reinterpret_cast<const double2*>(p + 2 * q);
    r[2 * q] = v.x;
    r[2 * q + 1] = v.y;
  }
}

constexpr int TP = 68;  // LDS row pitch (elements): 16-byte aligned rows, conflict-free 16-byte reads

template <typename T>
__device__ __forceinline__ typename MF<T>::acc tile_product(const T* a, const T* __restrict__ brow) {
  T bb[16];
  load16(brow, bb);
  typename MF<T>::acc c = {0, 0, 0, 0};
#pragma unroll
  for (int ks = 0; ks < 16; ++ks) c = MF<T>::step(a[ks], bb[ks], c);
  return c;
}

template <typename T>
__global__ __launch_bounds__(256) void kron64_pass_a(const T* __restrict__ K2, const T* __restrict__ K3,
                                                     const T* __restrict__ X, int64_t ldx, T* __restrict__ U,
                                                     int64_t ldu, const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ __align__(16) T Tt[64 * TP];
  const int i1 = blockIdx.x, b = blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, li = lane & 15;
  const T* S = X + (int64_t)b * ldx + (int64_t)i1 * 4096;
  T a[16];
  // T = S K3^T: T[r][c] = sum_k S[r][k] K3[c][k]
  load16(S + (16 * w + li) * 64 + 16 * g, a);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const auto c = tile_product<T>(a, K3 + (16 * t + li) * 64 + 16 * g);
    // c[it] = T[16w + row(g, it)][16t + li] -> Tt[16t + li][16w + row(g, it)]
    T* dst = Tt + (16 * t + li) * TP + 16 * w;
#pragma unroll
    for (int it = 0; it < 4; ++it) dst[MF<T>::row(g, it)] = c[it];
  }
  __syncthreads();
  // U = K2 T: U[r][c] = sum_k K2[r][k] Tt[c][k]
  load16(K2 + (16 * w + li) * 64 + 16 * g, a);
  T* Ub = U + (int64_t)b * ldu + (int64_t)i1 * 4096;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const auto c = tile_product<T>(a, Tt + (16 * t + li) * TP + 16 * g);
#pragma unroll
    for (int it = 0; it < 4; ++it) Ub[(16 * w + MF<T>::row(g, it)) * 64 + 16 * t + li] = c[it];
  }
}

// Y[b][j1][c] = d * scale * sum_k K1[j1][k] U[b][k][c] + shift * X[b][j1][c], c in tile
template <typename T>
__global__ __launch_bounds__(256) void kron64_pass_b(const T* __restrict__ K1, const T* __restrict__ U, int64_t ldu,
                                                     const T* __restrict__ X, int64_t ldx, T* __restrict__ Y,
                                                     int64_t ldy, T scale, T shift, const T* __restrict__ dvec,
                                                     float* __restrict__ partials, int nparts,
                                                     const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ __align__(16) T Ut[64 * TP];
  __shared__ float red[4];
  const int ct = blockIdx.x, b = blockIdx.y;
  const int c0 = 64 * ct;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, li = lane & 15;
  const T* Ub = U + (int64_t)b * ldu + c0;
  // stage the 64x64 tile transposed: Ut[c][k] = U[k][c0 + c]
  constexpr int VW = 16 / sizeof(T);
  for (int idx = threadIdx.x; idx < 64 * 64 / VW; idx += 256) {
    const int k = idx / (64 / VW), cc = (idx % (64 / VW)) * VW;
    T v[VW];
    if constexpr (VW == 4) {
      const float4 q = *reinterpret_cast<const float4*>(Ub + (int64_t)k * 4096 + cc);
      v[0] = q.x;
      v[1] = q.y;
      v[2] = q.z;
      v[3] = q.w;
    } else {
      const double2 q = *reinterpret_cast<const double2*>(Ub + (int64_t)k * 4096 + cc);
      v[0] = q.x;
      v[1] = q.y;
    }
#pragma unroll
    for (int e = 0; e < VW; ++e) Ut[(cc + e) * TP + k] = v[e];
  }
  __syncthreads();
  T a[16];
  load16(K1 + (16 * w + li) * 64 + 16 * g, a);
  float dot = 0.f;
  const int64_t xb = (int64_t)b * ldx, yb = (int64_t)b * ldy;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const auto c = tile_product<T>(a, Ut + (16 * t + li) * TP + 16 * g);
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int64_t off = (int64_t)(16 * w + MF<T>::row(g, it)) * 4096 + c0 + 16 * t + li;
      const T x = X[xb + off];
      T y = scale * c[it];
      if (dvec) y *= dvec[off];
      y += shift * x;
      Y[yb + off] = y;
      dot += (float)(x * y);
    }
  }
  if (partials) {
    const float s = ck::block_sum<256>(dot, red);
    if (threadIdx.x == 0) partials[(int64_t)b * nparts + ct] = s;
  }
}

template <typename T>
int launch_kron64(const T* const* f, int nrhs, const T* X, int64_t ldx, T* Y, int64_t ldy, T* U, T scale, T shift,
                  const T* dvec, float* partials, int nparts, const int* done, hipStream_t s) {
  const int64_t n = 64 * 64 * 64;
  hipLaunchKernelGGL(kron64_pass_a<T>, dim3(64, nrhs), dim3(256), 0, s, f[1], f[2], X, ldx, U, n, done);
  hipLaunchKernelGGL(kron64_pass_b<T>, dim3(64, nrhs), dim3(256), 0, s, f[0], (const T*)U, n, X, ldx, Y, ldy, scale,
                     shift, dvec, partials, nparts, done);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

}  // namespace

namespace ck {

bool kron64_eligible(const ck_op* op, const void* X, int64_t ldx, const void* Y, int64_t ldy) {
  if (op->nfactors != 3 || op->n != 64 * 64 * 64) return false;
  for (int i = 0; i < 3; ++i)
    if (op->fdim[i] != 64 || !op->factor[i] || !aligned16(op->factor[i])) return false;
  return aligned16(X) && aligned16(Y) && (ldx % 4 == 0) && (ldy % 4 == 0);
}

int kron64_apply_f32(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, void* work,
                     size_t work_bytes, const int* done, float* partials, int nparts, hipStream_t s) {
  if (!work || work_bytes < (size_t)nrhs * op->n * sizeof(float)) return CK_ERR_WORKSPACE;
  const float* f[3] = {op->factor[0], op->factor[1], op->factor[2]};
  return launch_kron64<float>(f, nrhs, X, ldx, Y, ldy, static_cast<float*>(work), op->scale, op->shift, nullptr,
                              partials, nparts, done, s);
}

// y = scale * Q (d * (Q^T x)) + shift * x, Q = Q0 x Q1 x Q2 (factor), Q^T factors in factor2, d = aux
int kron64_eig_apply_f32(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, void* work,
                         size_t work_bytes, const int* done, hipStream_t s) {
  const int64_t n = op->n;
  if (!work || work_bytes < 2 * (size_t)nrhs * n * sizeof(float)) return CK_ERR_WORKSPACE;
  for (int i = 0; i < 3; ++i)
    if (!op->factor2[i] || !aligned16(op->factor2[i])) return CK_ERR_BAD_ARG;
  if (!op->aux) return CK_ERR_BAD_ARG;
  float* U = static_cast<float*>(work);
  float* V = U + (size_t)nrhs * n;
  const float* ft[3] = {op->factor2[0], op->factor2[1], op->factor2[2]};
  const float* f[3] = {op->factor[0], op->factor[1], op->factor[2]};
  // V = d * (Q^T x)   (ldx of V = n)
  int st = launch_kron64<float>(ft, nrhs, X, ldx, V, n, U, 1.f, 0.f, op->aux, nullptr, 0, done, s);
  if (st) return st;
  // Y = scale * Q V + shift * X: pass B's shift term reads X at Y's layout, so apply it separately when needed
  if (op->shift == 0.f)
    return launch_kron64<float>(f, nrhs, V, n, Y, ldy, U, op->scale, 0.f, nullptr, nullptr, 0, done, s);
  return CK_ERR_UNSUPPORTED;
}

}  // namespace ck

extern "C" int ck_kron_apply_last2_f32(const float* K2, const float* K3, int32_t nrhs, const float* X, int64_t ldx,
                                       float* Y, int64_t ldy, void* stream) {
  const int64_t n = 64 * 64 * 64;
  if (!K2 || !K3 || !X || !Y || nrhs < 1 || ldx < n || ldy < n) return CK_ERR_BAD_ARG;
  if (!aligned16(K2) || !aligned16(K3) || !aligned16(X) || !aligned16(Y) || (ldx % 4) || (ldy % 4))
    return CK_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(kron64_pass_a<float>, dim3(64, nrhs), dim3(256), 0, (hipStream_t)stream, K2, K3, X, ldx, Y, ldy,
                     (const int*)nullptr);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

extern "C" int ck_kron_apply_f64(int32_t nfactors, const int32_t* fdim, const double* f0, const double* f1,
                                 const double* f2, double scale, double shift, int32_t nrhs, const double* X,
                                 int64_t ldx, double* Y, int64_t ldy, void* work, size_t work_bytes, void* stream) {
  const int64_t n = 64 * 64 * 64;
  if (nfactors != 3 || !fdim || !f0 || !f1 || !f2 || nrhs < 1 || !X || !Y || ldx < n || ldy < n) return CK_ERR_BAD_ARG;
  if (fdim[0] != 64 || fdim[1] != 64 || fdim[2] != 64) return CK_ERR_UNSUPPORTED;
  if (!aligned16(f0) || !aligned16(f1) || !aligned16(f2) || !aligned16(X) || !aligned16(Y) || (ldx % 2) || (ldy % 2))
    return CK_ERR_UNSUPPORTED;
  if (!work || work_bytes < (size_t)nrhs * n * sizeof(double)) return CK_ERR_WORKSPACE;
  const double* f[3] = {f0, f1, f2};
  return launch_kron64<double>(f, nrhs, X, ldx, Y, ldy, static_cast<double*>(work), scale, shift, nullptr, nullptr, 0,
                               nullptr, (hipStream_t)stream);
}
