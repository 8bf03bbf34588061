/*
 * cola_krylov.h -- C ABI of the MI355X (gfx950) Krylov engine behind cola_amd.
 *
 * Plain pointers and sizes only.  Every pointer argument named X, Y, V, x, r, p,
 * ... is DEVICE memory owned by the caller; the library never allocates or frees
 * user memory.  Every function returns an int status (CK_OK == 0) and never
 * throws.  `stream` is a hipStream_t passed as void*; all work is enqueued on it
 * and no function blocks the host unless its comment says so.
 *
 * Vector blocks ("column blocks"): B vectors of length n are stored one after
 * another, vector b at X + b*ldx (ldx >= n).  The reference keeps the same data
 * as an (n, B) row-major array (cola/ops/operator_base.py:56-60); the Python
 * layer transposes at the boundary.
 *
 * Reference citations are relative to the wilson-labs/cola source tree.
 */
#ifndef COLA_KRYLOV_H
#define COLA_KRYLOV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum {
  CK_OK = 0,
  CK_ERR_BAD_ARG = 1,     /* shapes, null pointers, sizes out of range          */
  CK_ERR_HIP = 2,         /* a HIP runtime call or launch failed                */
  CK_ERR_UNSUPPORTED = 3, /* operator structure this entry point does not cover */
  CK_ERR_WORKSPACE = 4,   /* workspace smaller than the *_workspace_bytes query */
  CK_ERR_NOCONV = 5,      /* tridiagonal QL did not converge in 60 sweeps/value  */
  /* Solver outcomes.  The enqueueing calls return before the solve runs, so these are
   * written by the kernels into the solver's control block (ck_cg_ctrl.status) and read
   * by the host with the step count; ck_cg_fused also returns them when it is asked to
   * wait for the result.  The reference records none of them: it stops on NaN (the loop
   * condition is false), divides safely on breakdown and returns at max_iters
   * (cola/linalg/inverse/cg.py:133-178), and so does the engine -- the status only says
   * which of these happened. */
  CK_ERR_MAXITER = 6,     /* the step cap was reached with a column above its tolerance */
  CK_ERR_BREAKDOWN = 7,   /* p.Ap <= 0 in an unfrozen column: A (or P) not positive definite */
  CK_ERR_NAN = 8,         /* a residual norm became NaN or infinite                     */
  CK_ERR_TIMEOUT = 9      /* a grid-wide wait of the persistent solve gave up (bounded spin) */
};

/* Library version string, e.g. "cola_krylov 0.1.0 gfx950". */
const char* ck_version(void);
/* Human-readable text for a status code. */
const char* ck_status_string(int status);

/* ---------------------------------------------------------------- operators
 * A square operator y = scale * S x + shift * x, where S is the structured part:
 *   CK_OP_IDENTITY   S = I
 *   CK_OP_DENSE_F32  S = A, row-major fp32 with leading dimension lda
 *                    (replaces Dense._matmat, cola/ops/operators.py:26-28)
 *   CK_OP_DENSE_BF16 S = A stored as bf16 (uint16 bit patterns), fp32 accumulate
 *   CK_OP_CSR_F32    S = CSR(rowptr, colind, values), int32 indices, nnz = rowptr[n]
 *                    stored entries (replaces Sparse._matmat, cola/ops/operators.py:64-78)
 *   CK_OP_KRON_F32   S = F0 (x) F1 (x) ... , 2 or 3 square row-major fp32 factors,
 *                    each of order <= 64 (replaces Kronecker._matmat,
 *                    cola/ops/operators.py:216-223)
 *   CK_OP_DIAG_F32   S = diag(a)  (cola/ops/operators.py:338-339)
 *   CK_OP_WOODBURY_INV  y = scale * (x - F (C^-1 (F^T x))): the inverse of
 *                    s I + F F^T when scale = 1/s and C^-1 = (s I + F^T F)^-1;
 *                    a = F (n x rank, row-major fp32, lda), aux = C^-1 (rank x rank
 *                    fp64); arithmetic after the loads in fp64, rank <= 128
 *   CK_OP_RBF_F32    S_ij = k(|p_i - p_j|) built on the fly from points a (n x dim fp32,
 *                    row-major, lda, dim <= 8) with lengthscale l = aux_f; rank selects
 *                    the stationary kernel: 0 exp(-r^2 / (2 l^2)), 1 exp(-r / l) (Matern
 *                    1/2), 2 (1 + a) exp(-a), a = sqrt(3) r / l (Matern 3/2), 3 (1 + a +
 *                    a^2/3) exp(-a), a = sqrt(5) r / l (Matern 5/2)
 *                    (replaces Kernel._matmat, cola/ops/operators.py:631-643)
 *   CK_OP_KRON_EIG_F32  S = Q diag(aux) Q^T with Q = factor[0] (x) factor[1] (x) factor[2]
 *                    (three 64x64 orthogonal factors) and factor2[i] = factor[i]^T:
 *                    a function of a Kronecker operator in its eigenbasis, e.g.
 *                    (K1 (x) K2 (x) K3 + s I)^-1 with aux = 1/(lambda + s); no engine
 *                    counterpart in the reference (it inverts bare Kronecker products
 *                    factor-wise, cola/linalg/inverse/inv.py:139-141); shift must be 0
 *   CK_OP_DENSE_BATCHED_F32  vector b is multiplied by its own row-major fp32 matrix
 *                    a + b * nnz (element stride nnz, leading dimension lda): many
 *                    independent operators of one order run as one launch (e.g. the
 *                    probes of a batched Lanczos run, each on its own instance); n,
 *                    lda, nnz, ldx, ldy multiples of 4, 16-byte aligned
 * The shift term is how Sum(S, ScalarMul(c) @ I) (cola/ops/operators.py:182-183,
 * 97-98) reaches one kernel epilogue.
 */
enum {
  CK_OP_IDENTITY = 0,
  CK_OP_DENSE_F32 = 1,
  CK_OP_DENSE_BF16 = 2,
  CK_OP_CSR_F32 = 3,
  CK_OP_KRON_F32 = 4,
  CK_OP_DIAG_F32 = 5,
  CK_OP_WOODBURY_INV = 6,
  CK_OP_RBF_F32 = 7,
  CK_OP_KRON_EIG_F32 = 8,
  CK_OP_DENSE_BATCHED_F32 = 9
};

typedef struct ck_op {
  int32_t kind;
  int32_t n;
  float scale;
  float shift;
  const void* a;
  int64_t lda;
  const int32_t* rowptr;
  const int32_t* colind;
  const float* values;
  int64_t nnz;
  int32_t nfactors;
  int32_t fdim[4];
  const float* factor[4];
  const float* factor2[4];
  const float* aux;
  int32_t rank;
  int32_t dim;
  float aux_f;
  int32_t flags;    /* CK_OP_SYMMETRIC: S = S^T is promised (dense fp32 of order a multiple of
                       128: the product reads only the blocks on and above the diagonal) */
  int32_t reserved[4];
} ck_op;
enum { CK_OP_SYMMETRIC = 1 };

/* Bytes of device scratch ck_op_apply needs for `nrhs` vectors. */
size_t ck_op_workspace_bytes(const ck_op* op, int32_t nrhs);

/* Y = op(X) for nrhs vectors.  X and Y must not overlap.
 * Replaces LinearOperator.__matmul__ -> _matmat (cola/ops/operator_base.py:97-106). */
int ck_op_apply(const ck_op* op, int32_t nrhs, const float* X, int64_t ldx, float* Y, int64_t ldy, void* work,
                size_t work_bytes, void* stream);

/* FP64 Kronecker matvec Y = scale * (F0 (x) F1 (x) F2) X + shift * X for three 64x64
 * row-major fp64 factors (v_mfma_f64_16x16x4f64); used for refinement residuals.
 * work >= nrhs * 64^3 doubles.  (cola/ops/operators.py:216-223 in FP64.) */
int ck_kron_apply_f64(int32_t nfactors, const int32_t* fdim, const double* f0, const double* f1, const double* f2,
                      double scale, double shift, int32_t nrhs, const double* X, int64_t ldx, double* Y, int64_t ldy,
                      void* work, size_t work_bytes, void* stream);
/* FP64 refinement residual of the 64^3 Kronecker operator in two passes: r = b - (scale
 * (f0 x f1 x f2) + shift I) x, written as fp32 R32, with norms[2c] = ||r_c||, norms[2c+1] =
 * ||b_c|| (fp64) -- ck_kron_apply_f64 followed by ck_ir_residual_f64 without storing A x.
 * work >= ck_kron_residual_workspace_bytes(nrhs). */
size_t ck_kron_residual_workspace_bytes(int32_t nrhs);
int ck_kron_residual_f64(const double* f0, const double* f1, const double* f2, double scale, double shift,
                         int32_t nrhs, const double* X, int64_t ldx, const double* Bv, int64_t ldb, float* R32,
                         int64_t ld32, double* norms, void* work, size_t work_bytes, void* stream);
/* One whole refinement step for a 64^3 Kronecker A (fp32 op, FP64 factor copies f0..f2 with
 * scale and shift) and its eigenbasis preconditioner P: z = P r32 and A z with their r.z and
 * z.A z partial sums, then, inside the FP64 residual's first pass, x64 = (accumulate ? x64 :
 * 0) + alpha z with alpha = (r.z)/(z.A z) per column (one PCG step from zero, as
 * ck_pcg_step1_f64), then r32 = b - A x64 and, when norms is not NULL, the norms of
 * ck_kron_residual_f64.  r32, z, Az: (nrhs, ld) fp32 blocks; x64, b64: fp64 with the same ld.
 * partials: ck_cg_partials_bytes(64^3, nrhs); work: A's and P's fp32 workspace; work64 >=
 * ck_kron_residual_workspace_bytes(nrhs).  No reference counterpart. */
int ck_ir_kron_step1_f64(const ck_op* A, const ck_op* P, const double* f0, const double* f1, const double* f2,
                         double scale, double shift, int32_t nrhs, float* r32, float* z, float* Az, int64_t ld,
                         double* x64, const double* b64, int32_t accumulate, double* norms, float* partials,
                         void* work, size_t work_bytes, void* work64, size_t work64_bytes, void* stream);

/* ---------------------------------------------------------------- mixed precision
 * Iterative refinement with FP64 residuals around fp32 inner solves (no
 * reference counterpart: cola's CG runs in the operator's dtype).
 * ck_ir_residual_f64: r = b - Ax (Ax may be NULL: r = b) in fp64, an optional fp32
 * copy r32, and norms[2*col] = ||r_col||, norms[2*col+1] = ||b_col|| (fp64).
 * ck_ir_update_f64: x += d.  ck_f32_to_bf16: round-to-nearest-even storage copy. */
size_t ck_ir_workspace_bytes(int32_t nrhs);
int ck_ir_residual_f64(int64_t n, int32_t nrhs, const double* b, int64_t ldb, const double* Ax, int64_t ldax,
                       double* r, int64_t ldr, float* r32, int64_t ldr32, double* norms, void* work,
                       size_t work_bytes, void* stream);
int ck_ir_update_f64(int64_t n, int32_t nrhs, double* x, int64_t ldx, const float* d, int64_t ldd, void* stream);
/* Start of a refinement solve: b64 = b, r32 = b, x64 = 0 (column blocks with leading
 * dimension ld) from b's element (i, c) at b[i*rs + c*cs], and norms as ck_ir_residual_f64
 * gives them for x = 0 (both ||b_c||).  x64 and norms may be NULL (not written; work is then
 * unused), else work >= ck_ir_workspace_bytes(nrhs). */
int ck_ir_load_rhs(int64_t n, int32_t nrhs, const float* b, int64_t rs, int64_t cs, double* b64, float* r32,
                   double* x64, int64_t ld, double* norms, void* work, size_t work_bytes, void* stream);
/* Inner-solve normalisation folded into the refinement (replaces ck_cg_normalize +
 * zeroing x + copying r, and ck_cg_finish + ck_ir_update_f64, inside a refinement step):
 * ck_ir_inner_start: r = r32 / max(||r||, 1e-40), x = 0, mult[col] = (float)||r|| with
 *   ||r|| = norms[2*col] from ck_ir_residual_f64; mult points at the CG state's
 *   CK_CG_MULT array, so the inner CG then runs from ck_cg_start as if normalised.
 * ck_ir_update_scaled_f64: x += (double)(d * mult[col]). */
int ck_ir_inner_start(int64_t n, int32_t nrhs, const float* r32, int64_t ld32, const double* norms, float* r,
                      float* x, int64_t ld, float* mult, void* stream);
int ck_ir_update_scaled_f64(int64_t n, int32_t nrhs, double* x, int64_t ldx, const float* d, int64_t ldd,
                            const float* mult, void* stream);
int ck_f32_to_bf16(int64_t count, const float* src, uint16_t* dst, void* stream);
/* dst = src for n floats (n % 4 == 0, 16-byte aligned): the bandwidth probe that sets the
 * measured HBM peak and calibrates byte counters; no reference counterpart. */
int ck_stream_copy_f32(int64_t n, const float* src, float* dst, void* stream);

/* Y = (I (x) K2 (x) K3) X for 64x64 fp32 factors: the fused mode-2/3 pass of the
 * 64^3 Kronecker matvec (one kernel; the per-mode loop of cola/ops/operators.py:218-222
 * for its last two modes). */
int ck_kron_apply_last2_f32(const float* K2, const float* K3, int32_t nrhs, const float* X, int64_t ldx, float* Y,
                            int64_t ldy, void* stream);

/* ---------------------------------------------------------------- vector kernels */
/* Per-column dot products out[b] = sum_i X[b,i] Y[b,i], fp64 result (a two-pass
 * reduction with fixed order).  Used for residual checks. */
int ck_dot_f32(int64_t n, int32_t nrhs, const float* X, int64_t ldx, const float* Y, int64_t ldy, double* out,
               void* work, size_t work_bytes, void* stream);
size_t ck_dot_workspace_bytes(int64_t n, int32_t nrhs);

/* Transpose between the (n, B) row-major layout and column blocks:
 * dst[b*ldd + i] = src[i*lds + b] (to_blocks=1) or dst[i*ldd + b] = src[b*lds + i]. */
int ck_transpose_f32(int64_t n, int32_t nrhs, const float* src, int64_t lds, float* dst, int64_t ldd, int32_t to_blocks,
                     void* stream);

/* ---------------------------------------------------------------- probes
 * Counter-based streams (SplitMix64 output function applied to (seed, stream, row)).
 * Rademacher probe p at row i is 1 - 2*bit(p % 64) of word(seed, p / 64, i).
 * Gaussian probe p at row i is Box-Muller of word(seed ^ 0x5851F42D4C957F2D, p, i).
 * Replaces the keyed torch.randn / sign draws of cola/linalg/trace/diagonal_estimation.py:187-190
 * and cola/linalg/tbd/slq.py:41. */
int ck_rng_rademacher_f32(int64_t n, int32_t nprobes, uint64_t seed, float* Z, int64_t ldz, void* stream);
int ck_rng_gaussian_f32(int64_t n, int32_t nprobes, uint64_t seed, float* Z, int64_t ldz, void* stream);
/* Packed Rademacher words: W[g*n + i] = word(seed, g, i) for g < ceil(nprobes/64). */
int ck_rng_rademacher_bits(int64_t n, int32_t ngroups, uint64_t seed, uint64_t* W, void* stream);

/* ---------------------------------------------------------------- CG
 * Batched (preconditioned) CG with the reference's arithmetic and stopping rule
 * (cola/linalg/inverse/cg.py:94-178): columns normalised by ||b||, per-column
 * threshold tol*||r0|| + tol, continue while ANY column has ||r|| > threshold and
 * k < max_iters, safe division at 1e-40, columns with ||r|| < 1e-40 frozen.
 * The loop condition is evaluated on the device; kernels of steps enqueued after
 * it turns false return immediately, so the host checks `done` only once per
 * chunk of steps.
 *
 * The state buffer (ck_cg_state_bytes) starts with this control block: */
typedef struct ck_cg_ctrl {
  int32_t k;         /* CG steps taken                                   */
  int32_t done;      /* 1 once the loop condition is false               */
  int32_t max_iters;
  int32_t nrhs;
  int32_t nparts;    /* partial sums per column (workgroups of the fused solve) */
  int32_t status;    /* CK_OK or CK_ERR_MAXITER / BREAKDOWN / NAN / TIMEOUT */
  int32_t pad[10];
} ck_cg_ctrl;
/* followed by fp32 arrays, each of ck_cg_stride(nrhs) floats, at these indices:
 * 0 mult(||b||) 1 tol 2 gamma 3 alpha 4 beta 5 rnorm, then errors[max_iters + 2]:
 * errors[k] = mean_b ||r_k|| (the value cola/utils/torch_tqdm.py:38-50 records). */
enum { CK_CG_MULT = 0, CK_CG_TOL = 1, CK_CG_GAMMA = 2, CK_CG_ALPHA = 3, CK_CG_BETA = 4, CK_CG_RNORM = 5, CK_CG_ERRORS = 6 };
int32_t ck_cg_stride(int32_t nrhs);
size_t ck_cg_state_bytes(int32_t nrhs, int32_t max_iters);
/* Device scratch for partial sums: 3 * nrhs * ck_cg_nparts(n) floats. */
int32_t ck_cg_nparts(int64_t n);
size_t ck_cg_partials_bytes(int64_t n, int32_t nrhs);

/* mult = ||b|| per column; bn = b / mult (safe division) -- cg.py:96-97. */
int ck_cg_normalize(int64_t n, int32_t nrhs, const float* b, int64_t ldb, float* bn, int64_t ldbn, void* state,
                    float* partials, void* stream);
/* p = z; gamma = r.z; rnorm = ||r||; tol = tol*rnorm + tol; errors[0]; k = 0;
 * done = !(any(rnorm > tol) && 0 < max_iters)  -- cg.py:98-101,122-138. */
int ck_cg_start(int64_t n, int32_t nrhs, const float* r, const float* z, float* p, int64_t ld, void* state, float tol,
                int32_t max_iters, float* partials, void* stream);
/* First half of take_cg_step (cg.py:141-149): alpha = gamma / p.Ap (safe, 0 when
 * frozen); x += alpha p; r -= alpha Ap. */
int ck_cg_step_xr(int64_t n, int32_t nrhs, const float* p, const float* Ap, float* x, float* r, int64_t ld, void* state,
                  float* partials, void* stream);
/* Second half (cg.py:150-154, 133-138): gamma1 = r.z, beta = gamma1/gamma (safe, 0
 * when frozen), p = z + beta p, k += 1, loop condition -> done.  z == r means no
 * preconditioner (r.r was already accumulated by ck_cg_step_xr). */
int ck_cg_step_p(int64_t n, int32_t nrhs, const float* r, const float* z, float* p, int64_t ld, void* state,
                 float* partials, void* stream);
/* Enqueue `nsteps` whole steps with library operators A and (optional) P.
 * z may equal r when P is NULL.  Ap and z are caller scratch blocks. */
int ck_cg_iterate(const ck_op* A, const ck_op* P, int32_t nrhs, float* x, float* r, float* p, float* Ap, float* z,
                  int64_t ld, void* state, float* partials, void* work, size_t work_bytes, int32_t nsteps,
                  void* stream);
/* The same nsteps steps, except that the last one stops after the x/r update: no
 * preconditioner application, p update or step scalars (k, done and errors keep the
 * previous step's values).  For callers that take x after exactly nsteps steps and
 * discard the Krylov state, e.g. the correction solves of iterative refinement
 * (no reference counterpart). */
int ck_cg_iterate_final(const ck_op* A, const ck_op* P, int32_t nrhs, float* x, float* r, float* p, float* Ap,
                        float* z, int64_t ld, void* state, float* partials, void* work, size_t work_bytes,
                        int32_t nsteps, void* stream);
/* ck_cg_iterate_final whose last step writes x nowhere but into an FP64 accumulator:
 * x64 += (double)((x + alpha p) * mult) with mult the state's CK_CG_MULT (the column
 * norms of the normalised start); x and r keep their previous values.  The refinement
 * update of a correction solve folded into its last CG step (no reference counterpart). */
int ck_cg_iterate_final_f64(const ck_op* A, const ck_op* P, int32_t nrhs, float* x, float* r, float* p, float* Ap,
                            float* z, int64_t ld, void* state, float* partials, void* work, size_t work_bytes,
                            int32_t nsteps, double* x64, int64_t ldx64, void* stream);
/* One preconditioned CG step from x = 0 on the residual block r, added to an FP64
 * accumulator: z = P r, alpha = (r.z)/(z.A z) per column (0 for a zero column),
 * x64 += (double)(alpha z).  The result of ck_cg_normalize + ck_cg_start +
 * ck_cg_iterate_final_f64(nsteps = 1) without the normalisation (one step is scale
 * invariant) in five launches: the preconditioner's passes emit r.z, A's emit z.A z.
 * Supported for a 64^3 Kronecker A with the Kronecker eigenbasis P; CK_ERR_UNSUPPORTED
 * otherwise.  partials: ck_cg_partials_bytes(n, nrhs); work: A's and P's workspace. */
int ck_pcg_step1_f64(const ck_op* A, const ck_op* P, int32_t nrhs, const float* r, float* z, float* Az, int64_t ld,
                     double* x64, int64_t ldx64, float* partials, void* work, size_t work_bytes, void* stream);
/* x *= mult; r *= mult  (cg.py:119). */
int ck_cg_finish(int64_t n, int32_t nrhs, float* x, float* r, int64_t ld, const void* state, void* stream);

/* ---------------------------------------------------------------- persistent fused CG
 * The whole batched CG solve of cola/linalg/inverse/cg.py:94-178 -- normalisation,
 * start, every step and the loop of cola/utils/torch_tqdm.py:86-90, un-normalisation --
 * in ONE cooperative launch (plus a one-block reset), for a dense fp32 operator
 * A = CK_OP_DENSE_F32 (y = scale*A x + shift*x) and P = NULL or CK_OP_DIAG_F32, 1 <= nrhs
 * <= 4, n*nrhs <= 7168 (and n <= 2560 for 3 or 4 columns): two n x nrhs fp32 vectors in LDS
 * and each thread's share of A p in registers.  Each workgroup keeps its rows of A in LDS
 * when they fit and the full r and p vectors; a step costs one grid-wide exchange: each
 * workgroup publishes its rows of A p as words tagged with the step, and every workgroup
 * reads all of A p back.  b, x0 (optional) and x, r (r optional) are (n x nrhs) arrays with
 * element strides (row stride, column stride), i.e. the reference's layout directly.
 * state: ck_cg_state_bytes(nrhs, max_iters), laid out and filled as by the chained
 * entry points (k, done, status, errors[0..k], per-column scalars); work:
 * ck_cg_fused_workspace_bytes.  nblocks = 0 picks the workgroup count.  Every grid-wide
 * wait gives up after spin_limit polling rounds (<= 0: about a second), records
 * CK_ERR_TIMEOUT in ctrl.status and ends the kernel; wait_offset (0 in normal use) is
 * added to the tag each wait expects, so a host can make every wait time out and exercise
 * that exit.  cooperative = 1 launches with hipLaunchCooperativeKernel (co-residency
 * guaranteed), 0 with a plain launch after the same occupancy check (about 20 us cheaper).
 * Returns CK_ERR_UNSUPPORTED for other operators or sizes. */
size_t ck_cg_fused_workspace_bytes(int64_t n, int32_t nrhs);
int ck_cg_fused_supported(const ck_op* A, const ck_op* P, int32_t nrhs);
int ck_cg_fused(const ck_op* A, const ck_op* P, int32_t nrhs, const float* b, int64_t b_rs, int64_t b_cs,
                const float* x0, int64_t x0_rs, int64_t x0_cs, float* x, int64_t x_rs, int64_t x_cs, float* r,
                int64_t r_rs, int64_t r_cs, void* state, float tol, int32_t max_iters, void* work,
                size_t work_bytes, int32_t nblocks, int64_t spin_limit, int32_t wait_offset, int32_t cooperative,
                void* stream);

/* ck_cg_fused without a workspace argument and without any allocation: the exchange buffer
 * is a file-scope device array of ck_cg_fused_noalloc_max_bytes() (64 KiB) per GPU, held
 * until the process ends.  No counterpart in the reference: for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * caller still passes b, x, r and state.  Sizes whose exchange buffer exceeds the array
 * (ck_cg_fused_workspace_bytes(n, nrhs) > ck_cg_fused_noalloc_max_bytes(), i.e. about
 * n * nrhs > 4096) return CK_ERR_WORKSPACE.  All calls on one GPU share the array: calls that overlap on one GPU
 * (other streams, other threads) race and must be serialised by the caller. */
int ck_cg_fused_noalloc(const ck_op* A, const ck_op* P, int32_t nrhs, const float* b, int64_t b_rs, int64_t b_cs,
                        const float* x0, int64_t x0_rs, int64_t x0_cs, float* x, int64_t x_rs, int64_t x_cs, float* r,
                        int64_t r_rs, int64_t r_cs, void* state, float tol, int32_t max_iters, int32_t nblocks,
                        int64_t spin_limit, int32_t wait_offset, int32_t cooperative, void* stream);
size_t ck_cg_fused_noalloc_max_bytes(void);

/* ---------------------------------------------------------------- Lanczos
 * Batched Lanczos with double classical Gram-Schmidt against every filled basis
 * vector (cola/linalg/decompositions/lanczos.py:235-296).  For nb start vectors
 * the basis is V[b][j][i] at V + b*ldb + j*ldv + i, j = 0..max_iters+1; column 0
 * is zero and column 1 the normalised start vector (lanczos.py:275-284).
 * diag (nb x max_iters) and subdiag (nb x (max_iters+1)) are fp32 like the
 * reference's alpha/beta arrays.  The state buffer starts with: */
typedef struct ck_lz_ctrl {
  int32_t i;         /* next step index (starts at 1)            */
  int32_t done;
  int32_t max_iters;
  int32_t nb;
  int32_t nparts;
  int32_t pad[11];
} ck_lz_ctrl;
/* followed by fp32 errors[max_iters + 2] (lanczos.py:256-261 values). */
size_t ck_lanczos_state_bytes(int32_t nb, int32_t max_iters);
size_t ck_lanczos_scratch_bytes(int64_t n, int32_t nb, int32_t max_iters);

/* V[:,0] = 0; V[:,1] = s / ||s||; diag = subdiag = 0; i = 1 (lanczos.py:275-284). */
int ck_lanczos_start(int64_t n, int32_t nb, const float* s, int64_t lds, float* V, int64_t ldv, int64_t ldb,
                     float* diag, float* subdiag, void* state, int32_t max_iters, void* scratch, void* stream);
/* Step `step` (0-based; uses column i = step+1), before the matvec:
 * V[:,i] /= ||V[:,i]|| (lanczos.py:240-241; for i > 1 the norm is subdiag[i-1]). */
int ck_lanczos_pre(int64_t n, int32_t nb, int32_t max_iters, float* V, int64_t ldv, int64_t ldb,
                   const float* subdiag, void* state, int32_t step, void* scratch, void* stream);
/* After the matvec wrote w = A v_i into column i+1: alpha_i = w.v_i; w -= alpha v_i +
 * beta_{i-1} v_{i-1}; w -= V(V^T w) twice over columns 1..i; beta_i = ||w||;
 * loop condition (lanczos.py:242-268) -> done. */
int ck_lanczos_post(int64_t n, int32_t nb, int32_t max_iters, float* V, int64_t ldv, int64_t ldb, float* diag,
                    float* subdiag, void* state, float tol, int32_t step, void* scratch, void* stream);
/* `nsteps` whole steps from `first_step` with library operator A. */
int ck_lanczos_iterate(const ck_op* A, int64_t n, int32_t nb, int32_t max_iters, float* V, int64_t ldv, int64_t ldb,
                       float* diag, float* subdiag, void* state, float tol, int32_t first_step, int32_t nsteps,
                       void* scratch, void* work, size_t work_bytes, void* stream);
/* ck_lanczos_iterate with block re-orthogonalisation (block = S in 2..16; no counterpart in
 * the reference, which re-orthogonalises every new vector against the whole basis,
 * lanczos.py:287-296): each step orthogonalises w (two classical passes) against its own
 * block of S columns and the column before it; when column i completes a block (i % S == 0,
 * i > S) the block [i-S+1, i] is made orthogonal to columns [1, i-S] by two passes of
 * H = V_old^T W, W -= V_old H on FP32 MFMA.  The recurrence, stop rule and outputs are those of
 * ck_lanczos_iterate.  ldv >= n rounded up to 64 (zero padding), 16-byte aligned V;
 * block_scratch >= ck_lanczos_block_scratch_bytes. */
size_t ck_lanczos_block_scratch_bytes(int64_t n, int32_t nb, int32_t block);
int ck_lanczos_iterate_block(const ck_op* A, int64_t n, int32_t nb, int32_t max_iters, float* V, int64_t ldv,
                             int64_t ldb, float* diag, float* subdiag, void* state, float tol, int32_t first_step,
                             int32_t nsteps, int32_t block, void* scratch, void* block_scratch,
                             size_t block_scratch_bytes, void* work, size_t work_bytes, void* stream);

/* ---------------------------------------------------------------- Arnoldi
 * Arnoldi with Gram-Schmidt against columns 0..idx (cola/linalg/decompositions/
 * arnoldi.py:289-324).  Q[b][j][i] at Q + b*ldb + j*ldv + i, j = 0..max_iters;
 * H[b][r][c] at H + b*(max_iters+1)*max_iters + r*max_iters + c (fp32).
 * mgs = 1 reproduces the reference's modified Gram-Schmidt order (one dependent
 * dot/axpy pair per column); mgs = 0 uses two classical passes (CGS2). */
typedef struct ck_ar_ctrl {
  int32_t idx;
  int32_t done;
  int32_t max_iters;
  int32_t nb;
  int32_t nparts;
  int32_t pad[11];
} ck_ar_ctrl;
size_t ck_arnoldi_state_bytes(int32_t nb, int32_t max_iters);
size_t ck_arnoldi_scratch_bytes(int64_t n, int32_t nb, int32_t max_iters);
int ck_arnoldi_start(int64_t n, int32_t nb, const float* s, int64_t lds, float* Q, int64_t ldv, int64_t ldb, float* H,
                     void* state, int32_t max_iters, void* scratch, void* stream);
/* After the matvec wrote w = A q_idx into column idx+1 (idx = step): Gram-Schmidt
 * against columns 0..idx into H[:, idx], H[idx+1, idx] = ||w||, w /= max(||w||, tol/2),
 * loop condition idx+1 < max_iters && any(||w|| > tol * H[1,0]) (arnoldi.py:293-319). */
int ck_arnoldi_post(int64_t n, int32_t nb, int32_t max_iters, float* Q, int64_t ldv, int64_t ldb, float* H,
                    void* state, float tol, int32_t step, int32_t mgs, void* scratch, void* stream);
int ck_arnoldi_iterate(const ck_op* A, int64_t n, int32_t nb, int32_t max_iters, float* Q, int64_t ldv, int64_t ldb,
                       float* H, void* state, float tol, int32_t first_step, int32_t nsteps, int32_t mgs,
                       void* scratch, void* work, size_t work_bytes, void* stream);

/* ---------------------------------------------------------------- small FP64 eigensolves
 * Symmetric tridiagonal eigensolve by implicit QL with Wilkinson shifts
 * (Bowdler, Martin, Reinsch & Wilkinson, Numer. Math. 11 (1968); EISPACK tql2),
 * one workgroup per problem.  Replaces xnp.eigh(T.to_dense()) of
 * lanczos.py:57, unary.py:52, slq.py:44, in FP64.
 * Problem b has diagonal d[b*ldd + 0..m-1] and off-diagonal e[b*lde + 0..m-2]
 * (fp32 inputs).  Outputs: ascending evals[b*m + c]; if evecs != NULL the
 * eigenvectors evecs[b*m*m + j*m + c] (component j of vector c); if first_row != NULL
 * the first components first_row[b*m + c]. status_out[b] = CK_OK or CK_ERR_NOCONV. */
int ck_tridiag_eigh_f64(int32_t nb, int32_t m, const float* d, int64_t ldd, const float* e, int64_t lde,
                        double* evals, double* evecs, double* first_row, int32_t* status_out, void* stream);
/* The same eigenproblems by Sturm multisection (L lanes per eigenvalue) and inverse
 * iteration (one thread per eigenvector; eigenvalues closer than 1e-3 ||T|| form a cluster
 * whose vectors are orthogonalised against each other), for lanczos.py:57 / unary.py:52 /
 * slq.py:44, where the QL recurrence's rotations are serial.  Same layouts as
 * ck_tridiag_eigh_f64, m <= 1024; first_row needs evecs; the first largest-magnitude
 * component of each vector is positive.  The inverse iteration keeps its factors in LDS when
 * the batch is small and in `work` (>= ck_tridiag_eigh_bisect_workspace_bytes(nb, m), may be
 * NULL: LDS then) for large batches.  status_out[b] = CK_OK or CK_ERR_NAN (a non-finite entry). */
size_t ck_tridiag_eigh_bisect_workspace_bytes(int32_t nb, int32_t m);
int ck_tridiag_eigh_bisect_f64(int32_t nb, int32_t m, const float* d, int64_t ldd, const float* e, int64_t lde,
                               double* evals, double* evecs, double* first_row, int32_t* status_out, void* work,
                               size_t work_bytes, void* stream);

/* out[c][i] = sum_j S[j][c] V[j][i] for c < k: Ritz vectors from a basis block
 * (lanczos.py:59, arnoldi.py:61).  V rows j = 0..m-1 at V + j*ldv; S (m x k) fp64
 * row-major with leading dimension lds; out vector c at out + c*ldo. */
int ck_basis_combine(int64_t n, int32_t m, int32_t k, const float* V, int64_t ldv, const double* S, int64_t lds,
                     float* out, int64_t ldo, void* stream);

/* out[b][i] = sum_j C[b][j] V[b][j][i] for j < k: one vector per probe from its own
 * basis block (f(A) z = Q_b c_b in cola/linalg/unary/unary.py:58-59).  C (nb x k) fp64. */
int ck_basis_combine_batched(int64_t n, int32_t nb, int32_t k, const float* V, int64_t ldv, int64_t ldb,
                             const double* C, int64_t ldc, float* out, int64_t ldo, void* stream);

/* ---------------------------------------------------------------- GMRES least squares
 * The small solve of GMRES (cola/linalg/inverse/gmres.py:99-121) for nb Arnoldi blocks H
 * ((m+1) x m fp32 row-major each, block b at H + b*ldh): y_b minimises
 * ||H_b[:rows, :cols] y - beta_b e1||^2 + sum_{padded i} y_i^2, then y_i = 0 on padded rows,
 * where a padded row's largest |entry| is below rtol * max|H_b| -- the normal equations
 * (H^T H + D) y = H^T beta e1 the reference solves (with rtol = 10 * tol), solved by Givens
 * rotations in FP64.  rows = cols + 1 after an Arnoldi run that stopped at step cols < m,
 * rows = cols = m after a full run; rows <= 4096.  y (nb x cols fp64, leading dimension ldy);
 * work >= ck_hessenberg_lstsq_workspace_bytes.  Replaces the xnp.solve of gmres.py:116. */
size_t ck_hessenberg_lstsq_workspace_bytes(int32_t nb, int32_t rows, int32_t cols);
int ck_hessenberg_lstsq_f64(int32_t nb, int32_t m, const float* H, int64_t ldh, int32_t rows, int32_t cols,
                            const float* beta, double rtol, void* work, size_t work_bytes, double* y, int64_t ldy,
                            void* stream);

/* Eigenpairs of nb small dense symmetric FP64 matrices (k <= 64), one workgroup each, by
 * parallel cyclic Jacobi: w (descending, stride_w apart) and E (row j = unit eigenvector of
 * w[j], leading dimension lde, stride_e apart).  Only the symmetric part of A is used.
 * sweeps[b] (optional) = sweeps taken, or -1 if max_sweeps ended before convergence.
 * Replaces the xnp.svd of the Nystrom factor (cola/linalg/preconditioning/
 * preconditioners.py:150), whose squared singular values and right vectors are the
 * eigenpairs of its Gram matrix. */
int ck_sym_eigh_f64(int32_t nb, int32_t k, const double* A, int64_t lda, int64_t stride_a, double* w,
                    int64_t stride_w, double* E, int64_t lde, int64_t stride_e, int32_t max_sweeps, int32_t* sweeps,
                    void* stream);

/* ---------------------------------------------------------------- sparse construction
 * COO -> CSR on the device (replaces the argsort + scipy coo_array(...).tocsr() of
 * Sparse.__init__, cola/ops/operators.py:64-75): rows in order, column indices ascending
 * inside each row, duplicate (row, col) entries summed in input order.  rows/cols are
 * int64 (nnz each); indptr (nrows + 1), indices and values (nnz each, an upper bound) are
 * written on the device, with the unique count in *nnz_out (device int32, optional).
 * *status (device int32) becomes 1 if any index lies outside [0, nrows) x [0, ncols);
 * those entries are dropped.  nrows, ncols, nnz < 2^31.  The call is asynchronous on
 * `stream`; work >= ck_coo_to_csr_workspace_bytes(nrows, nnz) (about 24 bytes/entry). */
size_t ck_coo_to_csr_workspace_bytes(int64_t nrows, int64_t nnz);
int ck_coo_to_csr_f32(int64_t nrows, int64_t ncols, int64_t nnz, const float* data, const int64_t* rows,
                      const int64_t* cols, int32_t* indptr, int32_t* indices, float* values, int32_t* nnz_out,
                      int32_t* status, void* work, size_t work_bytes, void* stream);
int ck_coo_to_csr_f64(int64_t nrows, int64_t ncols, int64_t nnz, const double* data, const int64_t* rows,
                      const int64_t* cols, int32_t* indptr, int32_t* indices, double* values, int32_t* nnz_out,
                      int32_t* status, void* work, size_t work_bytes, void* stream);
/* Y = scale A X + shift X with A a CSR matrix (int32 indices, fp32 values) and X, Y fp64
 * column blocks, FP64 accumulation: refinement residuals of sparse systems (Sparse._matmat,
 * cola/ops/operators.py:77-78, at higher precision). */
int ck_csr_apply_f64acc(int64_t n, const int32_t* rowptr, const int32_t* colind, const float* values, double scale,
                        double shift, int32_t nrhs, const double* X, int64_t ldx, double* Y, int64_t ldy,
                        void* stream);

/* ---------------------------------------------------------------- quadrature / trace
 * Stochastic Lanczos quadrature: per probe b, q_b = scale_b * sum_c tau_bc^2 f(theta_bc)
 * with f(theta) = 0 where |theta| <= 10*eps*max|theta| (cola/linalg/unary/unary.py:55-57,
 * cola/linalg/tbd/slq.py:46-51).  f_kind: 0 log, 1 identity, 2 exp, 3 sqrt, 4 reciprocal.
 * eps is the machine epsilon the threshold uses; max_mode 0 = max|theta|, 1 = max theta. */
int ck_slq_quadrature(int32_t nb, int32_t m, const double* evals, const double* first_row, const double* scale,
                      int32_t f_kind, double eps, int32_t max_mode, double* out, void* stream);

/* Hutchinson accumulation (cola/linalg/trace/diagonal_estimation.py:191-195, k = 0):
 * sum[i] += sum_b Z[b,i] AZ[b,i]; sumsq[i] += sum_b (Z[b,i] AZ[b,i])^2. */
int ck_hutch_accumulate(int64_t n, int32_t nprobes, const float* Z, int64_t ldz, const float* AZ, int64_t ldaz,
                        float* sum, float* sumsq, void* stream);
/* Fused CSR Hutchinson on packed Rademacher probes W (ck_rng_rademacher_bits):
 * one wavefront per row, lane = probe, so a probe block is never materialised;
 * the same sum/sumsq update as ck_hutch_accumulate for the 64*ngroups probes,
 * including the op's scale and shift. */
int ck_hutch_csr_bits(const ck_op* A, int32_t ngroups, const uint64_t* W, float* sum, float* sumsq, void* stream);
/* Trace-only CSR estimate from packed Rademacher words W (ck_rng_rademacher_bits):
 * out = scale * (1/(64*ngroups)) sum_i sum_k val_k * sum_g (64 - 2 popcount(W[g,i] ^ W[g,col_k]))
 *       + shift * n, fp64.  work >= ck_trace_workspace_bytes(n). */
int ck_trace_csr_rademacher_bits(const ck_op* A, int32_t ngroups, const uint64_t* W, double* out, void* work,
                                 size_t work_bytes, void* stream);
/* The same estimate with the words of ck_rng_rademacher_bits(seed) recomputed inside the
 * pass instead of read from memory: identical result, no probe buffer. */
int ck_trace_csr_rademacher_seeded(const ck_op* A, int32_t ngroups, uint64_t seed, double* out, void* work,
                                   size_t work_bytes, void* stream);
size_t ck_trace_workspace_bytes(int64_t n);

/* ---------------------------------------------------------------- learned preconditioner
 * No reference counterpart: the preconditioner these build enters the solver as
 * CG(P=...) (cola/linalg/inverse/cg.py:39-69, applied as z = P r at cg.py:150) in the
 * form P^-1 = (s I + F F^T) through CK_OP_WOODBURY_INV.
 *
 * Per-point MLP din -> 128 -> 128 -> dout with exact (erf) GELU after the two hidden
 * layers.  Parameters are one flat fp32 array: W1 (din x 128, row-major, y = x W + b),
 * b1, W2 (128 x 128), b2, W3 (128 x dout), b3.  din <= 8, dout <= 128.  params must
 * be 16-byte aligned (CK_ERR_BAD_ARG otherwise).
 * Points X are row-major (n x din, leading dimension ldx).  Output G[i*gsr + c*gsc].
 * With Z1, Z2 non-NULL the pre-activations are saved feature-major (Z[j*ldz + i]) for
 * ck_mlp_backward. */
int64_t ck_mlp_param_count(int32_t din, int32_t dout);
int ck_mlp_forward(int64_t n, int32_t din, int32_t dout, const float* X, int64_t ldx, const float* params, float* G,
                   int64_t gsr, int64_t gsc, float* Z1, float* Z2, int64_t ldz, void* stream);
/* grad = grad_scale * dL/dparams given dL/dG at dG[i*dsr + c*dsc] and the saved
 * pre-activations.  Deterministic: per-64-point partials summed in index order (FP64). */
size_t ck_mlp_grad_workspace_bytes(int64_t n, int32_t din, int32_t dout);
int ck_mlp_backward(int64_t n, int32_t din, int32_t dout, const float* X, int64_t ldx, const float* params,
                    const float* Z1, const float* Z2, int64_t ldz, const float* dG, int64_t dsr, int64_t dsc,
                    float grad_scale, float* grad, void* work, size_t work_bytes, void* stream);
/* Adam (Kingma & Ba 2015) in torch.optim.Adam's form, no weight decay; step >= 1. */
int ck_adam_step(int64_t count, float* params, const float* grad, float* m, float* v, float lr, float beta1,
                 float beta2, float eps, int32_t step, void* stream);
/* C (ka x kb, row-major fp64) = X^T Y for X[i*xsr + a*xsc], Y[i*ysr + b*ysc], i < n;
 * FP64 accumulation, 512-row partials summed in index order. */
size_t ck_gram_workspace_bytes(int64_t n, int32_t ka, int32_t kb);
int ck_gram_f64(int64_t n, int32_t ka, int32_t kb, const float* X, int64_t xsr, int64_t xsc, const float* Y,
                int64_t ysr, int64_t ysc, double* C, void* work, size_t work_bytes, void* stream);
/* Y[i, c] = alpha sum_j X[i, j] W[j][c] + beta Y[i, c] (beta == 0 ignores Y's old
 * contents); W (kin x kout) row-major fp32, kin <= 128. */
int ck_rows_gemm_f32(int64_t n, int32_t kin, int32_t kout, const float* X, int64_t xsr, int64_t xsc, const float* W,
                     float alpha, float beta, float* Y, int64_t ysr, int64_t ysc, void* stream);
/* Batched small SPD (k <= 64): A_b + (jitter_abs + jitter_rel * tr(A_b)/k) I = L L^T.
 * linv (batch x k x k, required) = L^-1; optional inv64/inv32 = A^-1, linvT32 = L^-T.
 * status[b] = 0, or j+1 when pivot j is not positive (outputs then undefined). */
int ck_spd_inverse_f64(int32_t batch, int32_t k, const double* A, int64_t lda, int64_t stride, double jitter_rel,
                       double jitter_abs, double* linv, double* inv64, float* inv32, float* linvT32, int32_t* status,
                       void* stream);
/* Captured-energy loss of features G against weighted target vectors U (w_i >= 0):
 * L = 1 - tr((G^T G)^-1 G^T U W U^T G) / sum w, given Sinv = (G^T G)^-1 (k x k) and
 * T = U^T G (m x k).  Writes C1 (m x k), C2 (k x k) with dL/dG = U C1 + G C2.  k, m <= 64. */
int ck_energy_loss_f64(int32_t k, int32_t m, const double* Sinv, const double* T, const double* w, float* C1,
                       float* C2, double* loss, void* stream);

/* Rank-r pivoted Cholesky of a dense symmetric PSD fp32 matrix (Harbrecht, Peters &
 * Schneider, Appl. Numer. Math. 62 (2012); the GPyTorch preconditioner of Gardner et al.
 * 2018): pivot j = argmax of the residual diagonal (ties: smallest index), column
 * L[j*ldl + i] = (A[p][i] - sum_{q<j} L_q[i] L_q[p]) / sqrt(d_p).  piv[j], dpiv[j] record
 * the pivots and their residual diagonals; a non-positive dpiv gives a zero column. */
size_t ck_pivchol_workspace_bytes(int64_t n);
int ck_pivchol_dense_f32(int64_t n, const float* A, int64_t lda, int32_t rank, float* L, int64_t ldl, int32_t* piv,
                         float* dpiv, void* work, size_t work_bytes, void* stream);
/* Dense RBF kernel matrix K_ij = exp(-|p_i - p_j|^2 / (2 l^2)) (the convention of
 * CK_OP_RBF_F32), exactly symmetric; K (fp32) and/or Kb (bf16, round to nearest even). */
int ck_rbf_dense_f32(int64_t n, int32_t dim, const float* P, int64_t ldp, float lengthscale, float* K, int64_t ldk,
                     uint16_t* Kb, int64_t ldkb, void* stream);
/* Y = scale A X + shift X with A fp32 (row-major) and X, Y fp64 column blocks, FP64
 * accumulation: refinement residuals of dense systems (Dense._matmat,
 * cola/ops/operators.py:26-28, at higher precision). */
int ck_dense_apply_f64acc(int64_t n, const float* A, int64_t lda, double scale, double shift, int32_t nrhs,
                          const double* X, int64_t ldx, double* Y, int64_t ldy, void* stream);

/* ---------------------------------------------------------------- BF16 MFMA sketch products
 * Yt[c][i] = scale * sum_l A[i][l] G[l][c] + shift * Gs[c][i], c < k (k = 32 or 64): the
 * sketch product K G that starts a low-rank preconditioner (the learned factor, a
 * Nystrom sketch: cola/linalg/preconditioning/preconditioners.py:147).  A is fp32 (n x n,
 * row-major, lda), read once and rounded to bf16 in registers; Gt is the sketch as k
 * feature-major bf16 rows (ldg), Gs its fp32 copy for the shift term; Yt fp32 feature-major.
 * v_mfma_f32_32x32x16_bf16 with FP32 accumulation; split = 1 rounds A to a bf16 pair
 * A_hi + A_lo (16 significant bits, two MFMAs per tile), split = 0 to one bf16 value.
 * ck_sketch_bf16_supported: n a multiple
 * of 256, lda % 4 == 0, ldg % 8 == 0, 16-byte aligned A and Gt. */
int ck_sketch_bf16_supported(int64_t n, int32_t k, int64_t lda, const void* A, int64_t ldg, const void* G);
int ck_sketch_bf16(int64_t n, int32_t k, const float* A, int64_t lda, float scale, float shift, const uint16_t* Gt,
                   int64_t ldg, const float* Gs, int64_t ldgs, float* Yt, int64_t ldy, int32_t split, void* stream);
/* Round rows x n fp32 values (row stride ldx) to bf16, round to nearest even: the bits to Xb
 * (ldb) and/or the rounded values back as fp32 to Xr (ldr); either output may be NULL. */
int ck_round_bf16(int64_t n, int32_t rows, const float* X, int64_t ldx, uint16_t* Xb, int64_t ldb, float* Xr,
                  int64_t ldr, void* stream);

/* ---------------------------------------------------------------- batched uneven systems
 * CG on many independent dense SPD systems (K_s + shift_s I) x_s = b_s of uneven order,
 * one rhs each: the blocks of a block-diagonal operator.  Replaces the per-block loop
 * inv(BlockDiag) = BlockDiag(inv(M_i)) (cola/linalg/inverse/inv.py:135-136, each block
 * then running cg.py:94-178 on its own) with one set of launches for all blocks.
 *
 * System s: order n[s]; matrix row-major at element offset moff[s] of A32 (fp32) and,
 * when prec[s] == 1, of A16 (bf16, same offsets), leading dimension lda[s] (multiple of
 * 8, padding columns zero); vectors at [voff[s], voff[s] + n[s]) of the concatenated
 * x, r, p, Ap.  Matvec tiles (tile_sys, tile_row0) cover rows_per_tile (4, 8, 16 or 32)
 * rows each; vector tiles (vtile_sys, vtile_i0) cover ck_bcg_vec_tile() entries; both
 * lists are ordered by system, with system s owning tiles [mt_begin[s], mt_end[s]) and
 * [vt_begin[s], vt_end[s]).  part_mv / part_vec hold one float per tile.  Per-system
 * state: gamma, alpha, beta, thresh (float), k, done (int32).  No reference counterpart
 * for the layout. */
typedef struct ck_bcg_plan {
  int32_t nsys;
  int32_t rows_per_tile;
  int32_t ntiles_mv;
  int32_t ntiles_vec;
  int32_t max_iters;
  int32_t pad0;
  const int32_t* tile_sys;
  const int32_t* tile_row0;
  const int32_t* vtile_sys;
  const int32_t* vtile_i0;
  const int32_t* n;
  const int32_t* lda;
  const int64_t* voff;
  const int64_t* moff;
  const int32_t* mt_begin;
  const int32_t* mt_end;
  const int32_t* vt_begin;
  const int32_t* vt_end;
  const float* shift;
  const int32_t* prec;
  const float* A32;
  const uint16_t* A16;
  float* x;
  float* r;
  float* p;
  float* Ap;
  float* part_mv;
  float* part_vec;
  float* gamma;
  float* alpha;
  float* beta;
  float* thresh;
  int32_t* k;
  int32_t* done;
  /* optional low-rank preconditioner P = (shift_s I + F_s F_s^T)^-1 per system (rank 0: none):
   * F_s row-major (n_s x rank) at element offset foff[s] of F; Cinv_s = (shift_s I + F_s^T F_s)^-1
   * (rank x rank, fp64) at Cinv + s*rank*rank; z, part_rz (ntiles_vec floats), part_ftr
   * (ntiles_vec * rank doubles) and u (nsys * rank doubles) are its scratch.  rank <= 64,
   * a divisor of 256. */
  int32_t rank;
  int32_t pad1;
  const float* F;
  const int64_t* foff;
  const double* Cinv;
  float* z;
  float* part_rz;
  double* part_ftr;
  double* u;
} ck_bcg_plan;

int32_t ck_bcg_vec_tile(void);
/* x = 0, r = p = b for systems with done[s] == 0 on entry (others are left alone);
 * thresh = tol_rel * ||b_s|| + tol_abs; systems already below it are marked done. */
int ck_bcg_start(const ck_bcg_plan* plan, const float* b, float tol_rel, float tol_abs, void* stream);
/* `iters` CG iterations for every system not done (five launches each, no host sync).
 * A system stops when ||r|| <= thresh or after max_iters iterations. */
int ck_bcg_iterate(const ck_bcg_plan* plan, int32_t iters, void* stream);
/* Ap = (K + shift I) p alone (tile-shape timing for the autotuner). */
int ck_bcg_matvec(const ck_bcg_plan* plan, void* stream);
/* Capture `iters` iterations on `stream` (not the legacy default stream) into a HIP
 * graph; *graph_out is an opaque handle for ck_graph_launch / ck_graph_destroy.  The
 * graph keeps the plan's pointers: the buffers must outlive it. */
int ck_bcg_graph_create(const ck_bcg_plan* plan, int32_t iters, void* stream, void** graph_out);
int ck_graph_launch(void* graph, void* stream);
int ck_graph_destroy(void* graph);
/* FP64 refinement on active systems (active[s] != 0): r32 = b - (K + shift I) x with K
 * as stored in A32, x and b fp64 and FP64 accumulation; rel[s] = ||r|| / ||b||.
 * part_rr, part_bb: ntiles_mv doubles each.  ck_bcg_update_f64: x64 += plan->x. */
int ck_bcg_residual_f64(const ck_bcg_plan* plan, const int32_t* active, const double* x64, const double* b64,
                        float* r32, double* part_rr, double* part_bb, double* rel, void* stream);
int ck_bcg_update_f64(const ck_bcg_plan* plan, const int32_t* active, double* x64, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* COLA_KRYLOV_H */
