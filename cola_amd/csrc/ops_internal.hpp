// Internal (non-ABI) entry points shared between the engine's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/cola_krylov.h"

namespace ck {

// op(X) -> Y; kernels return immediately when `done` (device flag) is nonzero.
int op_apply_internal(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, void* work,
                      size_t work_bytes, const int* done, hipStream_t s);
size_t op_workspace_internal(const ck_op* op, int nrhs);

// Kronecker apply that can also emit per-column partial sums of X*Y
// (nparts = kron_nparts(op) per column); used to fuse p.Ap into CG's matvec.
int kron_dispatch(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, void* work,
                  size_t work_bytes, const int* done, float* partials, int nparts, hipStream_t s);
int kron_nparts(const ck_op* op);
size_t kron_workspace_bytes(const ck_op* op, int nrhs);

bool kron64_eligible(const ck_op* op, const void* X, int64_t ldx, const void* Y, int64_t ldy);
int kron64_apply_f32(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, void* work,
                     size_t work_bytes, const int* done, float* partials, int nparts, hipStream_t s);
// rz_partials (optional): nrhs x 64 per-slab partial sums of X . Y (64^3 operators only)
int kron64_eig_apply_f32(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, void* work,
                         size_t work_bytes, const int* done, hipStream_t s, float* rz_partials = nullptr);

// Dense fp32 apply (3..8 rhs) that also writes per-block partial sums of X . Y:
// dense_xy_nparts(n) partials per rhs; CK_ERR_UNSUPPORTED outside that kernel's shapes.
inline int dense_xy_nparts(int64_t n) { return (int)((n + 31) / 32); }
inline size_t dense_xy_bytes(int64_t n, int nrhs) { return ((size_t)4 * nrhs * dense_xy_nparts(n) + 255) / 256 * 256; }
// With work (the operator's workspace), a CK_OP_SYMMETRIC operator of order a multiple of 128
// takes the half-reading symmetric product for 1..8 rhs.
int dense_apply_xy(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, const int* done,
                   float* xy_part, void* work, size_t work_bytes, hipStream_t s);

// CSR fp32 apply (average row length <= 32) that also writes per-block partial sums of X . Y:
// csr_xy_nparts(n) partials per rhs; CK_ERR_UNSUPPORTED for longer rows.
inline int csr_xy_nparts(int64_t n) { return (int)((n + 1023) / 1024); }
inline size_t csr_xy_bytes(int64_t n, int nrhs) { return ((size_t)4 * nrhs * csr_xy_nparts(n) + 255) / 256 * 256; }
int csr_apply_xy(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, const int* done,
                 float* xy_part, hipStream_t s);

int woodbury_apply(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, void* work,
                   size_t work_bytes, const int* done, hipStream_t s);
size_t woodbury_workspace_bytes(const ck_op* op, int nrhs);
// Scratch of CG's fused Woodbury step (rows per tile kWbFusedRows): p.Ap, r.r and r.z
// partials, per-tile F^T r partials and the core products u.
constexpr int kWbFusedRows = 128;
size_t wb_fused_bytes(int64_t n, int nrhs, int k);
// Offset of the symmetric dense product's scratch in an operator workspace: past the p.Ap
// partials and the fused Woodbury step's regions (rank <= 64) that CG keeps at the head.
inline size_t dense_sym_offset(int64_t n, int nrhs) {
  const size_t a = dense_xy_bytes(n, nrhs), b = wb_fused_bytes(n, nrhs, 64);
  return ((a > b ? a : b) + 255) / 256 * 256;
}
int rbf_apply(const ck_op* op, int nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, const int* done,
              hipStream_t s);

}  // namespace ck
