// Kronecker matvecs with 64x64 factors on MFMA (fp32: v_mfma_f32_16x16x4f32,
// fp64: v_mfma_f64_16x16x4f64).
//
// Layout: vector b of a column block is a 64x64x64 tensor X[i1][i2][i3].
//   pass A (per (b, i1) slab S = X[b][i1] as a 64x64 matrix; a block walks
//   several consecutive slabs):
//       U = K2 S K3^T          (modes 2 and 3 fused; the 64x64 intermediate
//                               stays in LDS, transposed)
//   pass B (per (b, 64-column tile of the i2*i3 plane); an fp32 block walks
//   four consecutive tiles):
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

// the four 16-column tiles of a 16x64 strip at once: four independent accumulation
// chains interleaved per k step keep the matrix pipe busy (one chain alone waits out
// each MFMA's latency before the next can issue)
template <typename T>
__device__ __forceinline__ void tile_product4(const T* a, const T* __restrict__ brow0, int64_t bstride,
                                              typename MF<T>::acc (&c)[4]) {
  T bb[4][16];
#pragma unroll
  for (int t = 0; t < 4; ++t) load16(brow0 + t * bstride, bb[t]);
#pragma unroll
  for (int t = 0; t < 4; ++t) c[t] = typename MF<T>::acc{0, 0, 0, 0};
#pragma unroll
  for (int ks = 0; ks < 16; ++ks)
#pragma unroll
    for (int t = 0; t < 4; ++t) c[t] = MF<T>::step(a[ks], bb[t][ks], c[t]);
}

// Pass A walks SLABS consecutive slabs per block (SLABS divides 64, so a block never
// straddles two right-hand sides and every block has a full set).  The factor operands
// every slab reuses (a lane's four K3 row slices and its K2 row slice) stay in registers;
// the next slab's rows are loaded while the current one is on the MFMA pipe.  The
// scheduling barriers pin the prefetch ahead of the MFMAs and the stores behind them:
// left to itself the compiler sinks the prefetch into the MFMA stream and then waits for
// it together with the current slab's rows.  fp32 alternates two LDS intermediates (one
// barrier per slab); fp64 (34 KB per intermediate) keeps one and a second barrier, reads
// K3 from LDS instead of registers and walks two slabs per block, so that two blocks (two
// waves per SIMD, 68 KB of LDS each) share a CU: 40.6 -> 35.3 us for the refinement pass.
template <typename T>
constexpr int slabs_per_block() {
  return sizeof(T) == 4 ? 4 : 2;
}

//
// With R given (fp32 only), each slab also emits partials[b * 64 + i1] = sum over the slab of
// R * U: the r.z partial sums of a preconditioned CG step when this pass writes z = P r.
//
// With upd.z given (fp64 refinement residuals), the input is first updated in place:
// X = (upd.accumulate ? X : 0) + (double)(alpha_b * z), alpha_b = (r.z)/(z.A z) of column b from
// the two partial-sum rows (a one-step PCG correction), and the pass runs on the new X.
struct XUpdate {
  const float* z;
  int64_t ldz;
  const float* prz;
  const float* pzaz;
  int np;
  int accumulate;
};

__device__ __forceinline__ double part_total(const float* __restrict__ part, int np, int b) {
  double acc = 0.0;
  for (int i = threadIdx.x & 63; i < np; i += 64) acc += (double)part[(int64_t)b * np + i];
  return ck::wave_sum(acc);
}

template <typename T, bool UPD = false>
__global__ __launch_bounds__(256) void kron64_pass_a(const T* __restrict__ K2, const T* __restrict__ K3,
                                                     const T* __restrict__ X, int64_t ldx, T* __restrict__ U,
                                                     int64_t ldu, const int* __restrict__ done,
                                                     const T* __restrict__ R, int64_t ldr,
                                                     float* __restrict__ partials, XUpdate upd) {
  if (done && *done) return;
  constexpr int SLABS = slabs_per_block<T>();
  constexpr int NBUF = sizeof(T) == 4 ? 2 : 1;
  __shared__ __align__(16) T Tt[NBUF][64 * TP];
  __shared__ float red[4];
  __shared__ float alpha_s;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, li = lane & 15;
  const int s0 = blockIdx.x * SLABS;
  const int64_t row_off = (16 * w + li) * 64 + 16 * g;
  auto slab = [&](int sidx) { return X + (int64_t)(sidx >> 6) * ldx + (int64_t)(sidx & 63) * 4096; };
  float alpha = 0.f;
  if constexpr (UPD) {
    // all SLABS slabs of a block belong to one column (SLABS divides 64)
    if (threadIdx.x < 64) {
      const double rz = part_total(upd.prz, upd.np, s0 >> 6);
      const double zaz = part_total(upd.pzaz, upd.np, s0 >> 6);
      if (threadIdx.x == 0) {
        double den = zaz;
        if (fabs(den) < (double)(float)1e-40) den = (double)(float)1e-40;
        alpha_s = (float)((double)(float)rz / den);
      }
    }
    __syncthreads();
    alpha = alpha_s;
  }
  // the slab rows of this lane, after the optional update (written back to X)
  auto fetch = [&](int sidx, T* dst) {
    T* xp = const_cast<T*>(slab(sidx)) + row_off;
    if constexpr (UPD) {
      const float* zp = upd.z + (int64_t)(sidx >> 6) * upd.ldz + (int64_t)(sidx & 63) * 4096 + row_off;
      T zv[16];
      float zf[16];
      load16(zp, zf);
      if (upd.accumulate) load16(xp, dst);
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        zv[q] = (T)(alpha * zf[q]);
        dst[q] = upd.accumulate ? dst[q] + zv[q] : zv[q];
      }
#pragma unroll
      for (int q = 0; q < 16; ++q) xp[q] = dst[q];
    } else {
      load16(xp, dst);
    }
  };
  // fp64 reads K3 from LDS: its 128 registers would leave the wave at the register limit
  // (an MFMA every 64 cycles leaves the LDS port idle anyway); fp32 keeps it in registers.
  // Row pitch 65 puts a wave's 16 rows in distinct bank pairs.
  constexpr bool K3L = sizeof(T) == 8;
  constexpr int K3P = 65;
  __shared__ T K3s[K3L ? 64 * K3P : 1];
  if constexpr (K3L) {
    T kv[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) kv[q] = K3[threadIdx.x + 256 * q];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int e = threadIdx.x + 256 * q;
      K3s[(e >> 6) * K3P + (e & 63)] = kv[q];
    }
  }
  T a[16];
  fetch(s0, a);
  T k3[K3L ? 1 : 4][16];
  if constexpr (!K3L) {
#pragma unroll
    for (int t = 0; t < 4; ++t) load16(K3 + (16 * t + li) * 64 + 16 * g, k3[t]);
  }
  T k2[16];
  load16(K2 + (16 * w + li) * 64 + 16 * g, k2);
  if constexpr (K3L) __syncthreads();  // K3s complete
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int it = 0; it < SLABS; ++it) {
    const int sidx = s0 + it;
    T an[16];
    if (it + 1 < SLABS) fetch(sidx + 1, an);
    __builtin_amdgcn_sched_barrier(0);
    // T = S K3^T: T[r][c] = sum_k S[r][k] K3[c][k]
    typename MF<T>::acc c[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) c[t] = typename MF<T>::acc{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < 16; ++ks)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if constexpr (K3L)
          c[t] = MF<T>::step(a[ks], K3s[(16 * t + li) * K3P + 16 * g + ks], c[t]);
        else
          c[t] = MF<T>::step(a[ks], k3[t][ks], c[t]);
      }
    T* tb = Tt[it % NBUF];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      // c[t][q] = T[16w + row(g, q)][16t + li] -> Tt[16t + li][16w + row(g, q)]
      T* dst = tb + (16 * t + li) * TP + 16 * w;
#pragma unroll
      for (int q = 0; q < 4; ++q) dst[MF<T>::row(g, q)] = c[t][q];
    }
    __syncthreads();
    // U = K2 T: U[r][c] = sum_k K2[r][k] Tt[c][k]
    tile_product4<T>(k2, tb + li * TP + 16 * g, 16 * TP, c);
    T* Ub = U + (int64_t)(sidx >> 6) * ldu + (int64_t)(sidx & 63) * 4096;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int q = 0; q < 4; ++q) Ub[(16 * w + MF<T>::row(g, q)) * 64 + 16 * t + li] = c[t][q];
    __builtin_amdgcn_sched_barrier(0);
    if (R) {
      const T* Rb = R + (int64_t)(sidx >> 6) * ldr + (int64_t)(sidx & 63) * 4096;
      float dot = 0.f;
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) dot += (float)(Rb[(16 * w + MF<T>::row(g, q)) * 64 + 16 * t + li] * c[t][q]);
      const float sum = ck::block_sum<256>(dot, red);
      if (threadIdx.x == 0) partials[sidx] = sum;
    }
    if (it + 1 < SLABS) {
#pragma unroll
      for (int q = 0; q < 16; ++q) a[q] = an[q];
    }
    if (NBUF == 1) __syncthreads();  // every wave is done reading tb before it is rewritten
  }
}

// Pass B for fp32: four consecutive 64-column tiles of one right-hand side per block.
// Staging map: thread (wave w, lane g*16 + li) loads U[16r + li][cs .. cs+3] for r = 0..3,
// cs = 4 * (4w + g); a wave's transposed LDS writes Ut[cs + e][16r + li] then fall in
// distinct banks.  The next tile's U is loaded into registers while the current tile is
// on the MFMA pipe, and the staging buffer alternates (one barrier per tile besides the
// partial sum's).  Per tile the arithmetic, its order and the partial are those of the
// one-tile kernel below.
constexpr int kTilesPerBlockB = 4;

__global__ __launch_bounds__(256) void kron64_pass_b4(const float* __restrict__ K1, const float* __restrict__ U,
                                                      int64_t ldu, const float* __restrict__ X, int64_t ldx,
                                                      float* __restrict__ Y, int64_t ldy, float scale, float shift,
                                                      const float* __restrict__ dvec, float* __restrict__ partials,
                                                      int nparts, const int* __restrict__ done) {
  if (done && *done) return;
  constexpr int TILES = kTilesPerBlockB;
  __shared__ __align__(16) float Ut[2][64 * TP];
  __shared__ float red[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, li = lane & 15;
  const int b = blockIdx.x / (64 / TILES), ct0 = (blockIdx.x % (64 / TILES)) * TILES;
  const float* Ubase = U + (int64_t)b * ldu;
  const int64_t xb = (int64_t)b * ldx, yb = (int64_t)b * ldy;
  const bool need_x = shift != 0.f || partials != nullptr;
  const int cs = 4 * (4 * w + g);
  float4 uq[4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
    uq[r] = *reinterpret_cast<const float4*>(Ubase + (int64_t)(16 * r + li) * 4096 + 64 * ct0 + cs);
  float a[16];
  load16(K1 + (16 * w + li) * 64 + 16 * g, a);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int it = 0; it < TILES; ++it) {
    const int ct = ct0 + it, c0 = 64 * ct;
    float* ut = Ut[it % 2];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = 16 * r + li;
      ut[(cs + 0) * TP + k] = uq[r].x;
      ut[(cs + 1) * TP + k] = uq[r].y;
      ut[(cs + 2) * TP + k] = uq[r].z;
      ut[(cs + 3) * TP + k] = uq[r].w;
    }
    if (it + 1 < TILES) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        uq[r] = *reinterpret_cast<const float4*>(Ubase + (int64_t)(16 * r + li) * 4096 + c0 + 64 + cs);
    }
    float xv[4][4];
    if (need_x) {
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) xv[t][q] = X[xb + (int64_t)(16 * w + 4 * g + q) * 4096 + c0 + 16 * t + li];
    }
    __builtin_amdgcn_sched_barrier(0);
    __syncthreads();
    floatx4 c[4];
    tile_product4<float>(a, ut + li * TP + 16 * g, 16 * TP, c);
    float dot = 0.f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t off = (int64_t)(16 * w + 4 * g + q) * 4096 + c0 + 16 * t + li;
        const float x = need_x ? xv[t][q] : 0.f;
        float y = scale * c[t][q];
        if (dvec) y *= dvec[off];
        y += shift * x;
        Y[yb + off] = y;
        dot += x * y;
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    if (partials) {
      const float sum = ck::block_sum<256>(dot, red);
      if (threadIdx.x == 0) partials[(int64_t)b * nparts + ct] = sum;
    }
  }
}

// The two mode-1 products of the eigenbasis preconditioner Q diag(d) Q^T, fused:
//   V[b][:, c] = scale * Q1 (d[:, c] * (Q1^T U[b][:, c]))   for each column c of a tile.
// A 64-column tile holds every i1 of its columns, so both products and the diagonal
// scaling run on one LDS-resident tile; the tile walk, staging map and prefetch are those
// of kron64_pass_b4.  Q1T is Q1^T stored explicitly (row-major).
__global__ __launch_bounds__(256) void kron64_pass_bb(const float* __restrict__ Q1T, const float* __restrict__ Q1,
                                                      const float* __restrict__ U, int64_t ldu,
                                                      float* __restrict__ V, int64_t ldv, float scale,
                                                      const float* __restrict__ dvec, const int* __restrict__ done) {
  if (done && *done) return;
  constexpr int TILES = kTilesPerBlockB;
  __shared__ __align__(16) float Ut[2][64 * TP];
  __shared__ __align__(16) float Wt[64 * TP];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, li = lane & 15;
  const int b = blockIdx.x / (64 / TILES), ct0 = (blockIdx.x % (64 / TILES)) * TILES;
  const float* Ubase = U + (int64_t)b * ldu;
  const int64_t vb = (int64_t)b * ldv;
  const int cs = 4 * (4 * w + g);
  float4 uq[4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
    uq[r] = *reinterpret_cast<const float4*>(Ubase + (int64_t)(16 * r + li) * 4096 + 64 * ct0 + cs);
  float a1[16], a2[16];
  load16(Q1T + (16 * w + li) * 64 + 16 * g, a1);
  load16(Q1 + (16 * w + li) * 64 + 16 * g, a2);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int it = 0; it < TILES; ++it) {
    const int c0 = 64 * (ct0 + it);
    float* ut = Ut[it % 2];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = 16 * r + li;
      ut[(cs + 0) * TP + k] = uq[r].x;
      ut[(cs + 1) * TP + k] = uq[r].y;
      ut[(cs + 2) * TP + k] = uq[r].z;
      ut[(cs + 3) * TP + k] = uq[r].w;
    }
    if (it + 1 < TILES) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        uq[r] = *reinterpret_cast<const float4*>(Ubase + (int64_t)(16 * r + li) * 4096 + c0 + 64 + cs);
    }
    float dv[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int q = 0; q < 4; ++q) dv[t][q] = dvec[(int64_t)(16 * w + 4 * g + q) * 4096 + c0 + 16 * t + li];
    __builtin_amdgcn_sched_barrier(0);
    __syncthreads();
    floatx4 c[4];
    // c[t][q] = (Q1^T U)[16w + 4g + q][c0 + 16t + li]
    tile_product4<float>(a1, ut + li * TP + 16 * g, 16 * TP, c);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      float* dst = Wt + (16 * t + li) * TP + 16 * w + 4 * g;
#pragma unroll
      for (int q = 0; q < 4; ++q) dst[q] = c[t][q] * dv[t][q];
    }
    __syncthreads();
    tile_product4<float>(a2, Wt + li * TP + 16 * g, 16 * TP, c);
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int q = 0; q < 4; ++q) V[vb + (int64_t)(16 * w + 4 * g + q) * 4096 + c0 + 16 * t + li] = scale * c[t][q];
    __builtin_amdgcn_sched_barrier(0);
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
  typename MF<T>::acc ct4[4];
  tile_product4<T>(a, Ut + li * TP + 16 * g, 16 * TP, ct4);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const auto& c = ct4[t];
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int64_t off = (int64_t)(16 * w + MF<T>::row(g, it)) * 4096 + c0 + 16 * t + li;
      const T x = X[xb + off];
      T y = scale * c[it];
      if (dvec) y *= dvec[off];
      y = fma(shift, x, y);  // explicit: kron64_resid_b repeats this arithmetic bit for bit
      Y[yb + off] = y;
      dot += (float)(x * y);
    }
  }
  if (partials) {
    const float s = ck::block_sum<256>(dot, red);
    if (threadIdx.x == 0) partials[(int64_t)b * nparts + ct] = s;
  }
}

// FP64 refinement residual in pass B's epilogue: per element ax = scale * (K1 U) + shift * x
// (the arithmetic of kron64_pass_b<double>), r = b - ax; writes r as fp32 and per-tile sums
// of r^2 and b^2 (rpart[(b * 64 + tile) * 2 + {0, 1}]) instead of storing A x.
__global__ __launch_bounds__(256) void kron64_resid_b(const double* __restrict__ K1, const double* __restrict__ U,
                                                      int64_t ldu, const double* __restrict__ X, int64_t ldx,
                                                      const double* __restrict__ Bv, int64_t ldb,
                                                      float* __restrict__ R32, int64_t ld32, double scale,
                                                      double shift, double* __restrict__ rpart) {
  __shared__ __align__(16) double Ut[64 * TP];
  __shared__ double red[4];
  const int ct = blockIdx.x, b = blockIdx.y;
  const int c0 = 64 * ct;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, li = lane & 15;
  const double* Ub = U + (int64_t)b * ldu + c0;
  for (int idx = threadIdx.x; idx < 64 * 32; idx += 256) {
    const int k = idx / 32, cc = (idx % 32) * 2;
    const double2 q = *reinterpret_cast<const double2*>(Ub + (int64_t)k * 4096 + cc);
    Ut[cc * TP + k] = q.x;
    Ut[(cc + 1) * TP + k] = q.y;
  }
  __syncthreads();
  double a[16];
  load16(K1 + (16 * w + li) * 64 + 16 * g, a);
  double rr = 0.0, bb = 0.0;
  const int64_t xb = (int64_t)b * ldx, bbase = (int64_t)b * ldb, rb = (int64_t)b * ld32;
  MF<double>::acc ct4[4];
  tile_product4<double>(a, Ut + li * TP + 16 * g, 16 * TP, ct4);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int64_t off = (int64_t)(16 * w + MF<double>::row(g, it)) * 4096 + c0 + 16 * t + li;
      double y = scale * ct4[t][it];
      y = fma(shift, X[xb + off], y);
      const double bv = Bv[bbase + off];
      const double rv = bv - y;
      R32[rb + off] = (float)rv;
      rr += rv * rv;
      bb += bv * bv;
    }
  }
  rr = ck::block_sum<256>(rr, red);
  bb = ck::block_sum<256>(bb, red);
  if (threadIdx.x == 0) {
    rpart[((int64_t)b * 64 + ct) * 2] = rr;
    rpart[((int64_t)b * 64 + ct) * 2 + 1] = bb;
  }
}

// norms[2c] = sqrt(sum of the 64 tile sums of r^2), norms[2c+1] likewise for b^2 (fixed order)
__global__ __launch_bounds__(64) void kron64_resid_norms(const double* __restrict__ rpart, double* __restrict__ norms) {
  const int col = blockIdx.x, which = blockIdx.y;
  double v = rpart[((int64_t)col * 64 + threadIdx.x) * 2 + which];
  v = ck::wave_sum(v);
  if (threadIdx.x == 0) norms[2 * col + which] = sqrt(v);
}

template <typename T>
int launch_kron64(const T* const* f, int nrhs, const T* X, int64_t ldx, T* Y, int64_t ldy, T* U, T scale, T shift,
                  const T* dvec, float* partials, int nparts, const int* done, hipStream_t s) {
  const int64_t n = 64 * 64 * 64;
  const unsigned grid_a = (unsigned)(64 / slabs_per_block<T>() * nrhs);
  hipLaunchKernelGGL(kron64_pass_a<T>, dim3(grid_a), dim3(256), 0, s, f[1], f[2], X, ldx, U, n, done,
                     (const T*)nullptr, (int64_t)0, (float*)nullptr, XUpdate{});
  if constexpr (sizeof(T) == 4) {
    if (nparts != 0 && nparts != 64) return CK_ERR_BAD_ARG;
    hipLaunchKernelGGL(kron64_pass_b4, dim3(64 / kTilesPerBlockB * nrhs), dim3(256), 0, s, f[0], (const T*)U, n, X,
                       ldx, Y, ldy, scale, shift, dvec, partials, nparts, done);
  } else {
    hipLaunchKernelGGL(kron64_pass_b<T>, dim3(64, nrhs), dim3(256), 0, s, f[0], (const T*)U, n, X, ldx, Y, ldy,
                       scale, shift, dvec, partials, nparts, done);
  }
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

// y = scale * Q (d * (Q^T x)), Q = Q0 x Q1 x Q2 (factor), Q^T factors in factor2, d = aux.
// Three launches: modes 2+3 of Q^T (pass A), both mode-1 products with the diagonal in
// between (pass BB), modes 2+3 of Q (pass A, writing Y).  With rz_partials, the last pass
// also emits the 64 per-slab partial sums of X . Y per column (r.z of a PCG step).
int kron64_eig_apply_f32(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, void* work,
                         size_t work_bytes, const int* done, hipStream_t s, float* rz_partials) {
  const int64_t n = op->n;
  if (!work || work_bytes < 2 * (size_t)nrhs * n * sizeof(float)) return CK_ERR_WORKSPACE;
  for (int i = 0; i < 3; ++i)
    if (!op->factor2[i] || !aligned16(op->factor2[i])) return CK_ERR_BAD_ARG;
  if (!op->aux) return CK_ERR_BAD_ARG;
  if (op->shift != 0.f) return CK_ERR_UNSUPPORTED;
  float* U = static_cast<float*>(work);
  float* V = U + (size_t)nrhs * n;
  const unsigned grid_a = (unsigned)(64 / slabs_per_block<float>() * nrhs);
  hipLaunchKernelGGL(kron64_pass_a<float>, dim3(grid_a), dim3(256), 0, s, op->factor2[1], op->factor2[2], X, ldx, U,
                     n, done, (const float*)nullptr, (int64_t)0, (float*)nullptr, XUpdate{});
  hipLaunchKernelGGL(kron64_pass_bb, dim3(64 / kTilesPerBlockB * nrhs), dim3(256), 0, s, op->factor2[0],
                     op->factor[0], (const float*)U, n, V, n, op->scale, op->aux, done);
  hipLaunchKernelGGL(kron64_pass_a<float>, dim3(grid_a), dim3(256), 0, s, op->factor[1], op->factor[2],
                     (const float*)V, n, Y, ldy, done, rz_partials ? X : nullptr, ldx, rz_partials, XUpdate{});
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // namespace ck

extern "C" int ck_kron_apply_last2_f32(const float* K2, const float* K3, int32_t nrhs, const float* X, int64_t ldx,
                                       float* Y, int64_t ldy, void* stream) {
  const int64_t n = 64 * 64 * 64;
  if (!K2 || !K3 || !X || !Y || nrhs < 1 || ldx < n || ldy < n) return CK_ERR_BAD_ARG;
  if (!aligned16(K2) || !aligned16(K3) || !aligned16(X) || !aligned16(Y) || (ldx % 4) || (ldy % 4))
    return CK_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(kron64_pass_a<float>, dim3((unsigned)(64 / slabs_per_block<float>() * nrhs)), dim3(256), 0,
                     (hipStream_t)stream, K2, K3, X, ldx, Y, ldy, (const int*)nullptr, (const float*)nullptr,
                     (int64_t)0, (float*)nullptr, XUpdate{});
  CK_CHECK_LAUNCH();
  return CK_OK;
}

extern "C" size_t ck_kron_residual_workspace_bytes(int32_t nrhs) {
  return sizeof(double) * ((size_t)nrhs * 64 * 64 * 64 + (size_t)nrhs * 64 * 2);
}

extern "C" int ck_kron_residual_f64(const double* f0, const double* f1, const double* f2, double scale, double shift,
                                    int32_t nrhs, const double* X, int64_t ldx, const double* Bv, int64_t ldb,
                                    float* R32, int64_t ld32, double* norms, void* work, size_t work_bytes,
                                    void* stream) {
  const int64_t n = 64 * 64 * 64;
  if (!f0 || !f1 || !f2 || nrhs < 1 || !X || !Bv || !R32 || !norms || ldx < n || ldb < n || ld32 < n)
    return CK_ERR_BAD_ARG;
  if (!aligned16(f0) || !aligned16(f1) || !aligned16(f2) || !aligned16(X) || !aligned16(Bv) || (ldx % 2) || (ldb % 2))
    return CK_ERR_UNSUPPORTED;
  if (!work || work_bytes < ck_kron_residual_workspace_bytes(nrhs)) return CK_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  double* U = static_cast<double*>(work);
  double* rpart = U + (size_t)nrhs * n;
  hipLaunchKernelGGL(kron64_pass_a<double>, dim3((unsigned)(64 / slabs_per_block<double>() * nrhs)), dim3(256), 0, s,
                     f1, f2, X, ldx, U, n, (const int*)nullptr, (const double*)nullptr, (int64_t)0, (float*)nullptr,
                     XUpdate{});
  hipLaunchKernelGGL(kron64_resid_b, dim3(64, nrhs), dim3(256), 0, s, f0, (const double*)U, n, X, ldx, Bv, ldb, R32,
                     ld32, scale, shift, rpart);
  hipLaunchKernelGGL(kron64_resid_norms, dim3(nrhs, 2), dim3(64), 0, s, (const double*)rpart, norms);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

extern "C" int ck_ir_kron_step1_f64(const ck_op* A, const ck_op* P, const double* f0, const double* f1,
                                    const double* f2, double scale, double shift, int32_t nrhs, float* r32, float* z,
                                    float* Az, int64_t ld, double* x64, const double* b64, int32_t accumulate,
                                    double* norms, float* partials, void* work, size_t work_bytes, void* work64,
                                    size_t work64_bytes, void* stream) {
  const int64_t n = 64 * 64 * 64;
  if (!A || !P || !f0 || !f1 || !f2 || nrhs < 1 || nrhs > 1024 || !r32 || !z || !Az || !x64 || !b64 || !partials ||
      ld < n)
    return CK_ERR_BAD_ARG;
  if (A->kind != CK_OP_KRON_F32 || P->kind != CK_OP_KRON_EIG_F32 || !ck::kron64_eligible(A, z, ld, Az, ld) ||
      !ck::kron64_eligible(P, r32, ld, z, ld))
    return CK_ERR_UNSUPPORTED;
  if (!aligned16(f0) || !aligned16(f1) || !aligned16(f2) || !aligned16(x64) || !aligned16(b64) || (ld % 2))
    return CK_ERR_UNSUPPORTED;
  if (!work64 || work64_bytes < ck_kron_residual_workspace_bytes(nrhs)) return CK_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  // partial-sum rows in the CG layout (64 per column at n = 64^3): z.Az in row 0, r.z in row 2
  float* pzaz = partials;
  float* prz = partials + 2 * (size_t)nrhs * 64;
  int st = ck::kron64_eig_apply_f32(P, nrhs, r32, ld, z, ld, work, work_bytes, nullptr, s, prz);
  if (st) return st;
  st = ck::kron_dispatch(A, nrhs, z, ld, Az, ld, work, work_bytes, nullptr, pzaz, 64, s);
  if (st) return st;
  double* U = static_cast<double*>(work64);
  double* rpart = U + (size_t)nrhs * n;
  XUpdate upd{z, ld, prz, pzaz, 64, accumulate ? 1 : 0};
  hipLaunchKernelGGL((kron64_pass_a<double, true>), dim3((unsigned)(64 / slabs_per_block<double>() * nrhs)), dim3(256),
                     0, s, f1, f2, (const double*)x64, ld, U, n, (const int*)nullptr, (const double*)nullptr,
                     (int64_t)0, (float*)nullptr, upd);
  hipLaunchKernelGGL(kron64_resid_b, dim3(64, nrhs), dim3(256), 0, s, f0, (const double*)U, n, (const double*)x64, ld,
                     b64, ld, r32, ld, scale, shift, rpart);
  if (norms) hipLaunchKernelGGL(kron64_resid_norms, dim3(nrhs, 2), dim3(64), 0, s, (const double*)rpart, norms);
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
