// COO -> CSR on the device: the construction of cola's Sparse operator.
//
// cola/ops/operators.py:64-75 argsorts the row indices and hands the triplets to scipy's
// coo_array(...).tocsr() on the host; the CSR it keeps has its rows in order, the column
// indices of each row ascending and duplicate (row, col) entries summed.  Here:
//   1. count:    entries per row (global atomics), out-of-range indices flagged;
//   2. scan:     row starts = exclusive prefix sum of the counts (work-efficient block scan,
//                recursive over block totals);
//   3. scatter:  each entry goes to row start + an atomic cursor, as a 64-bit key
//                (column << 32 | entry index), so the order inside a row is not yet fixed;
//   4. sort:     each row's keys ascending -- one wave ranks a row of <= 64 keys, one
//                workgroup runs a bitonic network in LDS for <= 4096 keys and over global
//                memory beyond; the entry index makes every key unique, so the result is
//                the same on every run whatever order the atomics gave;
//   5. dedup:    first-of-(row, col) flags, their exclusive scan = output positions, and
//                each group's values summed in input order (scipy's sum for <= 2
//                duplicates, bit for bit; for more it sums in an order of its own).
// Memory traffic is a few passes over the triplets; every phase is bandwidth-bound and
// none is a GEMM.
#include "common.hpp"

using namespace ck;

namespace {

constexpr int kNT = 256;
constexpr int kItems = 8;                    // scan elements per thread
constexpr int kScanTile = kNT * kItems;      // 2048
constexpr int kWaveRow = 64;                 // rows this short: one wave ranks them
constexpr int kLdsRow = 4096;                // rows this short: bitonic network in LDS

__global__ __launch_bounds__(kNT) void fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kNT) p[i] = v;
}

__global__ __launch_bounds__(kNT) void count_rows(int64_t nnz, int64_t nrows, int64_t ncols,
                                                  const int64_t* __restrict__ rows, const int64_t* __restrict__ cols,
                                                  int32_t* __restrict__ cnt, int32_t* __restrict__ status) {
  for (int64_t e = (int64_t)blockIdx.x * kNT + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * kNT) {
    const int64_t r = rows[e], c = cols[e];
    if (r < 0 || r >= nrows || c < 0 || c >= ncols) {
      atomicOr(status, 1);
      continue;
    }
    atomicAdd(cnt + r, 1);
  }
}

// exclusive scan of one tile of kScanTile elements; the tile's total goes to sums[tile]
__global__ __launch_bounds__(kNT) void scan_tiles(const int32_t* __restrict__ in, int64_t len,
                                                  int32_t* __restrict__ out, int32_t* __restrict__ sums) {
  __shared__ int32_t wsum[kNT / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kItems;
  int32_t v[kItems];
  int32_t t = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    v[i] = base + i < len ? in[base + i] : 0;
    t += v[i];
  }
  // inclusive scan of the thread totals across the wave, then across the four waves
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int32_t incl = t;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t u = __shfl_up(incl, o);
    if (lane >= o) incl += u;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int32_t before = 0;
  for (int i = 0; i < w; ++i) before += wsum[i];
  int32_t run = before + incl - t;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    if (base + i < len) out[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == kNT - 1) sums[blockIdx.x] = before + incl;
}

__global__ __launch_bounds__(kNT) void add_tile_offsets(int32_t* __restrict__ out, int64_t len,
                                                        const int32_t* __restrict__ offs) {
  const int32_t add = offs[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  for (int i = threadIdx.x; i < kScanTile; i += kNT)
    if (base + i < len) out[base + i] += add;
}

size_t scan_scratch_elems(int64_t len) {
  size_t total = 0;
  while (len > 1) {
    len = (len + kScanTile - 1) / kScanTile;
    total += (size_t)len + 1;
  }
  return total + 1;
}

// out[0..len) = exclusive scan of in[0..len); out may alias in.  scratch holds
// scan_scratch_elems(len) ints.
void exclusive_scan(const int32_t* in, int32_t* out, int64_t len, int32_t* scratch, hipStream_t s) {
  const int64_t tiles = (len + kScanTile - 1) / kScanTile;
  int32_t* sums = scratch;
  hipLaunchKernelGGL(scan_tiles, dim3((unsigned)tiles), dim3(kNT), 0, s, in, len, out, sums);
  if (tiles > 1) {
    exclusive_scan(sums, sums, tiles, scratch + tiles + 1, s);
    hipLaunchKernelGGL(add_tile_offsets, dim3((unsigned)tiles), dim3(kNT), 0, s, out, len, (const int32_t*)sums);
  }
}

__global__ __launch_bounds__(kNT) void scatter_keys(int64_t nnz, int64_t nrows, int64_t ncols,
                                                    const int64_t* __restrict__ rows, const int64_t* __restrict__ cols,
                                                    const int32_t* __restrict__ start, int32_t* __restrict__ cursor,
                                                    uint64_t* __restrict__ keys, int32_t* __restrict__ row_of) {
  for (int64_t e = (int64_t)blockIdx.x * kNT + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * kNT) {
    const int64_t r = rows[e], c = cols[e];
    if (r < 0 || r >= nrows || c < 0 || c >= ncols) continue;
    const int32_t pos = start[r] + atomicAdd(cursor + r, 1);
    keys[pos] = ((uint64_t)c << 32) | (uint64_t)e;
    row_of[pos] = (int32_t)r;
  }
}

// rows longer than one wave go on two lists, served by the workgroup sorters
__global__ __launch_bounds__(kNT) void classify_rows(int64_t nrows, const int32_t* __restrict__ start,
                                                     int32_t* __restrict__ lists, int32_t* __restrict__ counts) {
  for (int64_t r = (int64_t)blockIdx.x * kNT + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * kNT) {
    const int32_t len = start[r + 1] - start[r];
    if (len > kLdsRow) lists[nrows + atomicAdd(counts + 1, 1)] = (int32_t)r;
    else if (len > kWaveRow) lists[atomicAdd(counts, 1)] = (int32_t)r;
  }
}

// one wave per row of <= 64 keys: each lane's rank = number of smaller keys
__global__ __launch_bounds__(kNT) void sort_short_rows(int64_t nrows, const int32_t* __restrict__ start,
                                                       uint64_t* __restrict__ keys) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (kNT / 64);
  for (int64_t r = (int64_t)blockIdx.x * (kNT / 64) + (threadIdx.x >> 6); r < nrows; r += waves) {
    const int32_t s0 = start[r], len = start[r + 1] - s0;
    if (len < 2 || len > kWaveRow) continue;
    const uint64_t k = lane < len ? keys[s0 + lane] : ~0ull;
    int rank = 0;
    for (int j = 0; j < len; ++j) {
      const uint64_t kj = ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(k >> 32), j) << 32) |
                          (uint64_t)(uint32_t)__shfl((int)(uint32_t)k, j);
      rank += kj < k;
    }
    if (lane < len) keys[s0 + rank] = k;
  }
}

// ascending bitonic network on a[0..len) in the "flip" form (every compare puts the smaller
// key first), so positions >= len act as +infinity and are never touched
__device__ void bitonic_block(uint64_t* a, int64_t len) {
  int64_t P = 1;
  while (P < len) P <<= 1;
  const int64_t pairs = P >> 1;
  for (int64_t size = 2; size <= P; size <<= 1) {
    const int64_t half = size >> 1;
    for (int64_t t = threadIdx.x; t < pairs; t += kNT) {
      const int64_t blk = t / half, off = t % half;
      const int64_t i = blk * size + off, j = blk * size + size - 1 - off;
      if (j < len) {
        const uint64_t x = a[i], y = a[j];
        if (y < x) {
          a[i] = y;
          a[j] = x;
        }
      }
    }
    __syncthreads();
    for (int64_t stride = half >> 1; stride > 0; stride >>= 1) {
      for (int64_t t = threadIdx.x; t < pairs; t += kNT) {
        const int64_t i = (t / stride) * 2 * stride + t % stride, j = i + stride;
        if (j < len) {
          const uint64_t x = a[i], y = a[j];
          if (y < x) {
            a[i] = y;
            a[j] = x;
          }
        }
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(kNT) void sort_mid_rows(const int32_t* __restrict__ start,
                                                     const int32_t* __restrict__ list,
                                                     const int32_t* __restrict__ count, uint64_t* __restrict__ keys) {
  __shared__ uint64_t s[kLdsRow];
  const int32_t nlist = *count;
  for (int32_t i = blockIdx.x; i < nlist; i += gridDim.x) {
    const int32_t r = list[i];
    const int32_t s0 = start[r], len = start[r + 1] - s0;
    for (int t = threadIdx.x; t < len; t += kNT) s[t] = keys[s0 + t];
    __syncthreads();
    bitonic_block(s, len);
    for (int t = threadIdx.x; t < len; t += kNT) keys[s0 + t] = s[t];
    __syncthreads();
  }
}

__global__ __launch_bounds__(kNT) void sort_long_rows(int64_t nrows, const int32_t* __restrict__ start,
                                                      const int32_t* __restrict__ lists,
                                                      const int32_t* __restrict__ count, uint64_t* __restrict__ keys) {
  const int32_t nlist = *count;
  for (int32_t i = blockIdx.x; i < nlist; i += gridDim.x) {
    const int32_t r = lists[nrows + i];
    const int32_t s0 = start[r], len = start[r + 1] - s0;
    bitonic_block(keys + s0, len);
  }
}

__global__ __launch_bounds__(kNT) void first_flags(const int32_t* __restrict__ total, const int32_t* __restrict__ start,
                                                   const int32_t* __restrict__ row_of,
                                                   const uint64_t* __restrict__ keys, int32_t* __restrict__ flags) {
  const int32_t nv = *total;
  for (int64_t p = (int64_t)blockIdx.x * kNT + threadIdx.x; p < nv; p += (int64_t)gridDim.x * kNT) {
    const int32_t r = row_of[p];
    flags[p] = (p == start[r] || (keys[p - 1] >> 32) != (keys[p] >> 32)) ? 1 : 0;
  }
}

template <typename T>
__global__ __launch_bounds__(kNT) void emit_entries(const int32_t* __restrict__ total, const int32_t* __restrict__ start,
                                                    const int32_t* __restrict__ row_of,
                                                    const uint64_t* __restrict__ keys, const int32_t* __restrict__ outpos,
                                                    const T* __restrict__ data, int32_t* __restrict__ indices,
                                                    T* __restrict__ values) {
  const int32_t nv = *total;
  for (int64_t p = (int64_t)blockIdx.x * kNT + threadIdx.x; p < nv; p += (int64_t)gridDim.x * kNT) {
    const uint64_t k = keys[p];
    const int32_t r = row_of[p];
    if (!(p == start[r] || (keys[p - 1] >> 32) != (k >> 32))) continue;
    const int32_t end = start[r + 1];
    T v = data[(uint32_t)k];
    for (int32_t q = (int32_t)p + 1; q < end && (keys[q] >> 32) == (k >> 32); ++q) v += data[(uint32_t)keys[q]];
    const int32_t o = outpos[p];
    indices[o] = (int32_t)(k >> 32);
    values[o] = v;
  }
}

__global__ __launch_bounds__(kNT) void write_indptr(int64_t nrows, const int32_t* __restrict__ start,
                                                    const int32_t* __restrict__ outpos, int32_t* __restrict__ indptr,
                                                    int32_t* __restrict__ nnz_out) {
  for (int64_t r = (int64_t)blockIdx.x * kNT + threadIdx.x; r <= nrows; r += (int64_t)gridDim.x * kNT) {
    const int32_t v = outpos[start[r]];
    indptr[r] = v;
    if (r == nrows && nnz_out) *nnz_out = v;
  }
}

// Y = scale A X + shift X for CSR A (fp32 values) and FP64 column blocks X, Y, FP64
// accumulation: one wave per row, lanes strided over the row's entries
__global__ __launch_bounds__(kNT) void csr_apply_f64acc(int64_t n, const int32_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ colind,
                                                        const float* __restrict__ values, double scale, double shift,
                                                        int nrhs, const double* __restrict__ X, int64_t ldx,
                                                        double* __restrict__ Y, int64_t ldy) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (kNT / 64);
  for (int64_t r = (int64_t)blockIdx.x * (kNT / 64) + (threadIdx.x >> 6); r < n; r += waves) {
    const int32_t s0 = rowptr[r], s1 = rowptr[r + 1];
    for (int b = 0; b < nrhs; ++b) {
      const double* xb = X + (int64_t)b * ldx;
      double acc = 0.0;
      for (int32_t p = s0 + lane; p < s1; p += 64) acc = fma((double)values[p], xb[colind[p]], acc);
      acc = wave_sum(acc);
      if (lane == 0) Y[(int64_t)b * ldy + r] = scale * acc + shift * xb[r];
    }
  }
}

struct CooWork {
  int32_t *cnt, *start, *row_of, *flags, *lists, *counts, *scratch;
  uint64_t* keys;
};

size_t align_up(size_t v) { return (v + 255) / 256 * 256; }

size_t carve(int64_t nrows, int64_t nnz, void* base, CooWork* w) {
  const size_t scan_len = (size_t)(nrows + 1 > nnz + 1 ? nrows + 1 : nnz + 1);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += align_up(bytes);
    return at;
  };
  const size_t o_keys = take(8 * (size_t)nnz);
  const size_t o_cnt = take(4 * (size_t)(nrows + 1));
  const size_t o_start = take(4 * (size_t)(nrows + 1));
  const size_t o_row = take(4 * (size_t)nnz);
  const size_t o_flags = take(4 * (size_t)(nnz + 1));
  const size_t o_lists = take(4 * (size_t)(2 * nrows));
  const size_t o_counts = take(4 * 4);
  const size_t o_scr = take(4 * scan_scratch_elems((int64_t)scan_len));
  if (w && base) {
    char* b = static_cast<char*>(base);
    w->keys = reinterpret_cast<uint64_t*>(b + o_keys);
    w->cnt = reinterpret_cast<int32_t*>(b + o_cnt);
    w->start = reinterpret_cast<int32_t*>(b + o_start);
    w->row_of = reinterpret_cast<int32_t*>(b + o_row);
    w->flags = reinterpret_cast<int32_t*>(b + o_flags);
    w->lists = reinterpret_cast<int32_t*>(b + o_lists);
    w->counts = reinterpret_cast<int32_t*>(b + o_counts);
    w->scratch = reinterpret_cast<int32_t*>(b + o_scr);
  }
  return off;
}

unsigned grid_for(int64_t n) {
  const int64_t g = (n + kNT - 1) / kNT;
  return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

template <typename T>
int coo_to_csr(int64_t nrows, int64_t ncols, int64_t nnz, const T* data, const int64_t* rows, const int64_t* cols,
               int32_t* indptr, int32_t* indices, T* values, int32_t* nnz_out, int32_t* status, void* work,
               size_t work_bytes, hipStream_t s) {
  if (nrows < 0 || ncols < 0 || nnz < 0 || nrows >= INT32_MAX || ncols >= INT32_MAX || nnz >= INT32_MAX || !indptr ||
      !status || (nnz > 0 && (!data || !rows || !cols || !indices || !values)))
    return CK_ERR_BAD_ARG;
  if (!work || work_bytes < carve(nrows, nnz, nullptr, nullptr)) return CK_ERR_WORKSPACE;
  CooWork w;
  carve(nrows, nnz, work, &w);
  hipLaunchKernelGGL(fill_i32, dim3(1), dim3(kNT), 0, s, status, (int64_t)1, 0);
  hipLaunchKernelGGL(fill_i32, dim3(grid_for(nrows + 1)), dim3(kNT), 0, s, w.cnt, nrows + 1, 0);
  hipLaunchKernelGGL(fill_i32, dim3(1), dim3(kNT), 0, s, w.counts, (int64_t)4, 0);
  if (nnz > 0)
    hipLaunchKernelGGL(count_rows, dim3(grid_for(nnz)), dim3(kNT), 0, s, nnz, nrows, ncols, rows, cols, w.cnt, status);
  // start[0..nrows] = exclusive scan of cnt (cnt[nrows] = 0, so start[nrows] = valid entries)
  exclusive_scan(w.cnt, w.start, nrows + 1, w.scratch, s);
  if (nnz > 0) {
    hipLaunchKernelGGL(fill_i32, dim3(grid_for(nrows + 1)), dim3(kNT), 0, s, w.cnt, nrows + 1, 0);
    hipLaunchKernelGGL(scatter_keys, dim3(grid_for(nnz)), dim3(kNT), 0, s, nnz, nrows, ncols, rows, cols,
                       (const int32_t*)w.start, w.cnt, w.keys, w.row_of);
    if (nrows > 0) {
      hipLaunchKernelGGL(classify_rows, dim3(grid_for(nrows)), dim3(kNT), 0, s, nrows, (const int32_t*)w.start,
                         w.lists, w.counts);
      const int64_t wg = (nrows + kNT / 64 - 1) / (kNT / 64);
      hipLaunchKernelGGL(sort_short_rows, dim3((unsigned)(wg > 65536 ? 65536 : wg)), dim3(kNT), 0, s, nrows,
                         (const int32_t*)w.start, w.keys);
      hipLaunchKernelGGL(sort_mid_rows, dim3(1024), dim3(kNT), 0, s, (const int32_t*)w.start,
                         (const int32_t*)w.lists, (const int32_t*)w.counts, w.keys);
      hipLaunchKernelGGL(sort_long_rows, dim3(256), dim3(kNT), 0, s, nrows, (const int32_t*)w.start,
                         (const int32_t*)w.lists, (const int32_t*)(w.counts + 1), w.keys);
    }
    hipLaunchKernelGGL(fill_i32, dim3(grid_for(nnz + 1)), dim3(kNT), 0, s, w.flags, nnz + 1, 0);
    hipLaunchKernelGGL(first_flags, dim3(grid_for(nnz)), dim3(kNT), 0, s, (const int32_t*)(w.start + nrows),
                       (const int32_t*)w.start, (const int32_t*)w.row_of, (const uint64_t*)w.keys, w.flags);
  } else {
    hipLaunchKernelGGL(fill_i32, dim3(1), dim3(kNT), 0, s, w.flags, (int64_t)1, 0);
  }
  // flags[p] -> output position; flags[valid] = 0 so outpos[valid] = unique count
  exclusive_scan(w.flags, w.flags, nnz + 1, w.scratch, s);
  if (nnz > 0)
    hipLaunchKernelGGL(emit_entries<T>, dim3(grid_for(nnz)), dim3(kNT), 0, s, (const int32_t*)(w.start + nrows),
                       (const int32_t*)w.start, (const int32_t*)w.row_of, (const uint64_t*)w.keys,
                       (const int32_t*)w.flags, data, indices, values);
  hipLaunchKernelGGL(write_indptr, dim3(grid_for(nrows + 1)), dim3(kNT), 0, s, nrows, (const int32_t*)w.start,
                     (const int32_t*)w.flags, indptr, nnz_out);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // namespace

extern "C" {

size_t ck_coo_to_csr_workspace_bytes(int64_t nrows, int64_t nnz) {
  if (nrows < 0 || nnz < 0) return 0;
  return carve(nrows, nnz, nullptr, nullptr);
}

int ck_coo_to_csr_f32(int64_t nrows, int64_t ncols, int64_t nnz, const float* data, const int64_t* rows,
                      const int64_t* cols, int32_t* indptr, int32_t* indices, float* values, int32_t* nnz_out,
                      int32_t* status, void* work, size_t work_bytes, void* stream) {
  return coo_to_csr<float>(nrows, ncols, nnz, data, rows, cols, indptr, indices, values, nnz_out, status, work,
                           work_bytes, (hipStream_t)stream);
}

int ck_coo_to_csr_f64(int64_t nrows, int64_t ncols, int64_t nnz, const double* data, const int64_t* rows,
                      const int64_t* cols, int32_t* indptr, int32_t* indices, double* values, int32_t* nnz_out,
                      int32_t* status, void* work, size_t work_bytes, void* stream) {
  return coo_to_csr<double>(nrows, ncols, nnz, data, rows, cols, indptr, indices, values, nnz_out, status, work,
                            work_bytes, (hipStream_t)stream);
}

int ck_csr_apply_f64acc(int64_t n, const int32_t* rowptr, const int32_t* colind, const float* values, double scale,
                        double shift, int32_t nrhs, const double* X, int64_t ldx, double* Y, int64_t ldy,
                        void* stream) {
  if (n < 1 || nrhs < 1 || !rowptr || !X || !Y || ldx < n || ldy < n) return CK_ERR_BAD_ARG;
  const int64_t wg = (n + kNT / 64 - 1) / (kNT / 64);
  hipLaunchKernelGGL(csr_apply_f64acc, dim3((unsigned)(wg > 65536 ? 65536 : wg)), dim3(kNT), 0,
                     (hipStream_t)stream, n, rowptr, colind, values, scale, shift, (int)nrhs, X, ldx, Y, ldy);
  CK_CHECK_LAUNCH();
  return CK_OK;
}

}  // extern "C"
