"""ctypes binding of libcolakrylov.so (the engine's C ABI, include/cola_krylov.h).

The library is built in-tree by ``cola_amd.csrc.build`` (``__graft_entry__.build()``).
Loading fails loudly: every engine entry point needs it, and nothing in
``cola_amd`` substitutes host code for a missing kernel.
"""
import ctypes
import os

import torch

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libcolakrylov.so")

(CK_OK, CK_ERR_BAD_ARG, CK_ERR_HIP, CK_ERR_UNSUPPORTED, CK_ERR_WORKSPACE, CK_ERR_NOCONV, CK_ERR_MAXITER,
 CK_ERR_BREAKDOWN, CK_ERR_NAN, CK_ERR_TIMEOUT) = range(10)
(CK_OP_IDENTITY, CK_OP_DENSE_F32, CK_OP_DENSE_BF16, CK_OP_CSR_F32, CK_OP_KRON_F32, CK_OP_DIAG_F32,
 CK_OP_WOODBURY_INV, CK_OP_RBF_F32, CK_OP_KRON_EIG_F32, CK_OP_DENSE_BATCHED_F32) = range(10)
CK_OP_SYMMETRIC = 1  # ck_op.flags
CK_CG_MULT, CK_CG_TOL, CK_CG_GAMMA, CK_CG_ALPHA, CK_CG_BETA, CK_CG_RNORM, CK_CG_ERRORS = range(7)
CTRL_BYTES = 64


class KrylovError(RuntimeError):
    """A C-ABI call returned a nonzero status."""

    def __init__(self, fn, status, msg):
        super().__init__(f"{fn} failed with status {status}: {msg}")
        self.status = status


class UnsupportedOperator(KrylovError):
    pass


class CkOp(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("scale", ctypes.c_float),
        ("shift", ctypes.c_float),
        ("a", ctypes.c_void_p),
        ("lda", ctypes.c_int64),
        ("rowptr", ctypes.c_void_p),
        ("colind", ctypes.c_void_p),
        ("values", ctypes.c_void_p),
        ("nnz", ctypes.c_int64),
        ("nfactors", ctypes.c_int32),
        ("fdim", ctypes.c_int32 * 4),
        ("factor", ctypes.c_void_p * 4),
        ("factor2", ctypes.c_void_p * 4),
        ("aux", ctypes.c_void_p),
        ("rank", ctypes.c_int32),
        ("dim", ctypes.c_int32),
        ("aux_f", ctypes.c_float),
        ("flags", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 4),
    ]


class CkBcgPlan(ctypes.Structure):
    """Mirror of ck_bcg_plan in include/cola_krylov.h (batched uneven systems)."""
    _fields_ = [("nsys", ctypes.c_int32), ("rows_per_tile", ctypes.c_int32), ("ntiles_mv", ctypes.c_int32),
                ("ntiles_vec", ctypes.c_int32), ("max_iters", ctypes.c_int32), ("pad0", ctypes.c_int32)] + \
               [(name, ctypes.c_void_p) for name in
                ("tile_sys", "tile_row0", "vtile_sys", "vtile_i0", "n", "lda", "voff", "moff", "mt_begin", "mt_end",
                 "vt_begin", "vt_end", "shift", "prec", "A32", "A16", "x", "r", "p", "Ap", "part_mv", "part_vec",
                 "gamma", "alpha", "beta", "thresh", "k", "done")] + \
               [("rank", ctypes.c_int32), ("pad1", ctypes.c_int32)] + \
               [(name, ctypes.c_void_p) for name in ("F", "foff", "Cinv", "z", "part_rz", "part_ftr", "u")]


_i32, _i64, _u64, _f32, _f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_float, ctypes.c_double
_p, _sz = ctypes.c_void_p, ctypes.c_size_t
_op = ctypes.POINTER(CkOp)

# name: (restype, argtypes)
_PROTOS = {
    "ck_version": (ctypes.c_char_p, []),
    "ck_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "ck_op_workspace_bytes": (_sz, [_op, _i32]),
    "ck_op_apply": (ctypes.c_int, [_op, _i32, _p, _i64, _p, _i64, _p, _sz, _p]),
    "ck_kron_apply_f64": (ctypes.c_int, [_i32, _p, _p, _p, _p, _f64, _f64, _i32, _p, _i64, _p, _i64, _p, _sz, _p]),
    "ck_kron_apply_last2_f32": (ctypes.c_int, [_p, _p, _i32, _p, _i64, _p, _i64, _p]),
    "ck_ir_workspace_bytes": (_sz, [_i32]),
    "ck_ir_residual_f64": (ctypes.c_int, [_i64, _i32, _p, _i64, _p, _i64, _p, _i64, _p, _i64, _p, _p, _sz, _p]),
    "ck_ir_kron_step1_f64": (ctypes.c_int, [_op, _op, _p, _p, _p, ctypes.c_double, ctypes.c_double, _i32, _p, _p, _p,
                                            _i64, _p, _p, _i32, _p, _p, _p, _sz, _p, _sz, _p]),
    "ck_kron_residual_workspace_bytes": (ctypes.c_size_t, [_i32]),
    "ck_kron_residual_f64": (ctypes.c_int, [_p, _p, _p, ctypes.c_double, ctypes.c_double, _i32, _p, _i64, _p, _i64, _p,
                                            _i64, _p, _p, _sz, _p]),
    "ck_ir_update_f64": (ctypes.c_int, [_i64, _i32, _p, _i64, _p, _i64, _p]),
    "ck_ir_load_rhs": (ctypes.c_int, [_i64, _i32, _p, _i64, _i64, _p, _p, _p, _i64, _p, _p, _sz, _p]),
    "ck_ir_inner_start": (ctypes.c_int, [_i64, _i32, _p, _i64, _p, _p, _p, _i64, _p, _p]),
    "ck_ir_update_scaled_f64": (ctypes.c_int, [_i64, _i32, _p, _i64, _p, _i64, _p, _p]),
    "ck_f32_to_bf16": (ctypes.c_int, [_i64, _p, _p, _p]),
    "ck_stream_copy_f32": (ctypes.c_int, [_i64, _p, _p, _p]),
    "ck_dot_f32": (ctypes.c_int, [_i64, _i32, _p, _i64, _p, _i64, _p, _p, _sz, _p]),
    "ck_dot_workspace_bytes": (_sz, [_i64, _i32]),
    "ck_transpose_f32": (ctypes.c_int, [_i64, _i32, _p, _i64, _p, _i64, _i32, _p]),
    "ck_rng_rademacher_f32": (ctypes.c_int, [_i64, _i32, _u64, _p, _i64, _p]),
    "ck_rng_gaussian_f32": (ctypes.c_int, [_i64, _i32, _u64, _p, _i64, _p]),
    "ck_rng_rademacher_bits": (ctypes.c_int, [_i64, _i32, _u64, _p, _p]),
    "ck_cg_stride": (_i32, [_i32]),
    "ck_cg_state_bytes": (_sz, [_i32, _i32]),
    "ck_cg_nparts": (_i32, [_i64]),
    "ck_cg_partials_bytes": (_sz, [_i64, _i32]),
    "ck_cg_normalize": (ctypes.c_int, [_i64, _i32, _p, _i64, _p, _i64, _p, _p, _p]),
    "ck_cg_start": (ctypes.c_int, [_i64, _i32, _p, _p, _p, _i64, _p, _f32, _i32, _p, _p]),
    "ck_cg_step_xr": (ctypes.c_int, [_i64, _i32, _p, _p, _p, _p, _i64, _p, _p, _p]),
    "ck_cg_step_p": (ctypes.c_int, [_i64, _i32, _p, _p, _p, _i64, _p, _p, _p]),
    "ck_cg_iterate": (ctypes.c_int, [_op, _op, _i32, _p, _p, _p, _p, _p, _i64, _p, _p, _p, _sz, _i32, _p]),
    "ck_cg_iterate_final": (ctypes.c_int, [_op, _op, _i32, _p, _p, _p, _p, _p, _i64, _p, _p, _p, _sz, _i32, _p]),
    "ck_pcg_step1_f64": (ctypes.c_int, [_op, _op, _i32, _p, _p, _p, _i64, _p, _i64, _p, _p, _sz, _p]),
    "ck_cg_iterate_final_f64": (ctypes.c_int, [_op, _op, _i32, _p, _p, _p, _p, _p, _i64, _p, _p, _p, _sz, _i32, _p,
                                               _i64, _p]),
    "ck_cg_finish": (ctypes.c_int, [_i64, _i32, _p, _p, _i64, _p, _p]),
    "ck_hessenberg_lstsq_workspace_bytes": (_sz, [_i32, _i32, _i32]),
    "ck_hessenberg_lstsq_f64": (ctypes.c_int, [_i32, _i32, _p, _i64, _i32, _i32, _p, _f64, _p, _sz, _p, _i64, _p]),
    "ck_sym_eigh_f64": (ctypes.c_int, [_i32, _i32, _p, _i64, _i64, _p, _i64, _p, _i64, _i64, _i32, _p, _p]),
    "ck_coo_to_csr_workspace_bytes": (_sz, [_i64, _i64]),
    "ck_coo_to_csr_f32": (ctypes.c_int, [_i64, _i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _sz, _p]),
    "ck_coo_to_csr_f64": (ctypes.c_int, [_i64, _i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _sz, _p]),
    "ck_csr_apply_f64acc": (ctypes.c_int, [_i64, _p, _p, _p, _f64, _f64, _i32, _p, _i64, _p, _i64, _p]),
    "ck_lanczos_block_scratch_bytes": (_sz, [_i64, _i32, _i32]),
    "ck_lanczos_iterate_block": (ctypes.c_int, [_op, _i64, _i32, _i32, _p, _i64, _i64, _p, _p, _p, _f32, _i32, _i32,
                                                _i32, _p, _p, _sz, _p, _sz, _p]),
    "ck_sketch_bf16_supported": (ctypes.c_int, [_i64, _i32, _i64, _p, _i64, _p]),
    "ck_sketch_bf16": (ctypes.c_int, [_i64, _i32, _p, _i64, _f32, _f32, _p, _i64, _p, _i64, _p, _i64, _i32, _p]),
    "ck_round_bf16": (ctypes.c_int, [_i64, _i32, _p, _i64, _p, _i64, _p, _i64, _p]),
    "ck_cg_fused_workspace_bytes": (_sz, [_i64, _i32]),
    "ck_cg_fused_supported": (ctypes.c_int, [_op, _op, _i32]),
    "ck_cg_fused_noalloc": (ctypes.c_int, [_op, _op, _i32, _p, _i64, _i64, _p, _i64, _i64, _p, _i64, _i64, _p, _i64,
                                           _i64, _p, _f32, _i32, _i32, _i64, _i32, _i32, _p]),
    "ck_cg_fused_noalloc_max_bytes": (_sz, []),
    "ck_cg_fused": (ctypes.c_int, [_op, _op, _i32, _p, _i64, _i64, _p, _i64, _i64, _p, _i64, _i64, _p, _i64, _i64, _p,
                                   _f32, _i32, _p, _sz, _i32, _i64, _i32, _i32, _p]),
    "ck_lanczos_state_bytes": (_sz, [_i32, _i32]),
    "ck_lanczos_scratch_bytes": (_sz, [_i64, _i32, _i32]),
    "ck_lanczos_start": (ctypes.c_int, [_i64, _i32, _p, _i64, _p, _i64, _i64, _p, _p, _p, _i32, _p, _p]),
    "ck_lanczos_pre": (ctypes.c_int, [_i64, _i32, _i32, _p, _i64, _i64, _p, _p, _i32, _p, _p]),
    "ck_lanczos_post": (ctypes.c_int, [_i64, _i32, _i32, _p, _i64, _i64, _p, _p, _p, _f32, _i32, _p, _p]),
    "ck_lanczos_iterate": (ctypes.c_int, [_op, _i64, _i32, _i32, _p, _i64, _i64, _p, _p, _p, _f32, _i32, _i32, _p,
                                          _p, _sz, _p]),
    "ck_arnoldi_state_bytes": (_sz, [_i32, _i32]),
    "ck_arnoldi_scratch_bytes": (_sz, [_i64, _i32, _i32]),
    "ck_arnoldi_start": (ctypes.c_int, [_i64, _i32, _p, _i64, _p, _i64, _i64, _p, _p, _i32, _p, _p]),
    "ck_arnoldi_post": (ctypes.c_int, [_i64, _i32, _i32, _p, _i64, _i64, _p, _p, _f32, _i32, _i32, _p, _p]),
    "ck_arnoldi_iterate": (ctypes.c_int, [_op, _i64, _i32, _i32, _p, _i64, _i64, _p, _p, _f32, _i32, _i32, _i32, _p,
                                          _p, _sz, _p]),
    "ck_tridiag_eigh_f64": (ctypes.c_int, [_i32, _i32, _p, _i64, _p, _i64, _p, _p, _p, _p, _p]),
    "ck_tridiag_eigh_bisect_workspace_bytes": (_sz, [_i32, _i32]),
    "ck_tridiag_eigh_bisect_f64": (ctypes.c_int, [_i32, _i32, _p, _i64, _p, _i64, _p, _p, _p, _p, _p, _sz, _p]),
    "ck_basis_combine": (ctypes.c_int, [_i64, _i32, _i32, _p, _i64, _p, _i64, _p, _i64, _p]),
    "ck_basis_combine_batched": (ctypes.c_int, [_i64, _i32, _i32, _p, _i64, _i64, _p, _i64, _p, _i64, _p]),
    "ck_slq_quadrature": (ctypes.c_int, [_i32, _i32, _p, _p, _p, _i32, _f64, _i32, _p, _p]),
    "ck_hutch_accumulate": (ctypes.c_int, [_i64, _i32, _p, _i64, _p, _i64, _p, _p, _p]),
    "ck_hutch_csr_bits": (ctypes.c_int, [_op, _i32, _p, _p, _p, _p]),
    "ck_trace_csr_rademacher_bits": (ctypes.c_int, [_op, _i32, _p, _p, _p, _sz, _p]),
    "ck_trace_csr_rademacher_seeded": (ctypes.c_int, [_op, _i32, ctypes.c_uint64, _p, _p, _sz, _p]),
    "ck_trace_workspace_bytes": (_sz, [_i64]),
    # learned preconditioner, pivoted Cholesky, dense helpers
    "ck_mlp_param_count": (_i64, [_i32, _i32]),
    "ck_mlp_forward": (ctypes.c_int, [_i64, _i32, _i32, _p, _i64, _p, _p, _i64, _i64, _p, _p, _i64, _p]),
    "ck_mlp_grad_workspace_bytes": (_sz, [_i64, _i32, _i32]),
    "ck_mlp_backward": (ctypes.c_int, [_i64, _i32, _i32, _p, _i64, _p, _p, _p, _i64, _p, _i64, _i64, _f32, _p, _p,
                                       _sz, _p]),
    "ck_adam_step": (ctypes.c_int, [_i64, _p, _p, _p, _p, _f32, _f32, _f32, _f32, _i32, _p]),
    "ck_gram_workspace_bytes": (_sz, [_i64, _i32, _i32]),
    "ck_gram_f64": (ctypes.c_int, [_i64, _i32, _i32, _p, _i64, _i64, _p, _i64, _i64, _p, _p, _sz, _p]),
    "ck_rows_gemm_f32": (ctypes.c_int, [_i64, _i32, _i32, _p, _i64, _i64, _p, _f32, _f32, _p, _i64, _i64, _p]),
    "ck_spd_inverse_f64": (ctypes.c_int, [_i32, _i32, _p, _i64, _i64, _f64, _f64, _p, _p, _p, _p, _p, _p]),
    "ck_energy_loss_f64": (ctypes.c_int, [_i32, _i32, _p, _p, _p, _p, _p, _p, _p]),
    "ck_pivchol_workspace_bytes": (_sz, [_i64]),
    "ck_pivchol_dense_f32": (ctypes.c_int, [_i64, _p, _i64, _i32, _p, _i64, _p, _p, _p, _sz, _p]),
    "ck_rbf_dense_f32": (ctypes.c_int, [_i64, _i32, _p, _i64, _f32, _p, _i64, _p, _i64, _p]),
    "ck_dense_apply_f64acc": (ctypes.c_int, [_i64, _p, _i64, _f64, _f64, _i32, _p, _i64, _p, _i64, _p]),
    # batched uneven systems
    "ck_bcg_vec_tile": (_i32, []),
    "ck_bcg_start": (ctypes.c_int, [ctypes.POINTER(CkBcgPlan), _p, _f32, _f32, _p]),
    "ck_bcg_iterate": (ctypes.c_int, [ctypes.POINTER(CkBcgPlan), _i32, _p]),
    "ck_bcg_matvec": (ctypes.c_int, [ctypes.POINTER(CkBcgPlan), _p]),
    "ck_bcg_graph_create": (ctypes.c_int, [ctypes.POINTER(CkBcgPlan), _i32, _p, ctypes.POINTER(ctypes.c_void_p)]),
    "ck_graph_launch": (ctypes.c_int, [_p, _p]),
    "ck_graph_destroy": (ctypes.c_int, [_p]),
    "ck_bcg_residual_f64": (ctypes.c_int, [ctypes.POINTER(CkBcgPlan), _p, _p, _p, _p, _p, _p, _p, _p]),
    "ck_bcg_update_f64": (ctypes.c_int, [ctypes.POINTER(CkBcgPlan), _p, _p, _p]),
}

_lib = None


def lib_path():
    return _LIB_PATH


def load():
    """Load the library once; raise ImportError with the build command if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise ImportError(f"{_LIB_PATH} is missing: build it with `python -m cola_amd.csrc.build` "
                          "(or __graft_entry__.build()); the Krylov engine has no substitute for it")
    lib = ctypes.CDLL(_LIB_PATH)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            raise ImportError(f"{_LIB_PATH} does not export {name}: rebuild it with `python -m cola_amd.csrc.build`")
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def bind(name, restype, argtypes):
    """Bind an extra symbol (used by modules that add entry points)."""
    lib = load()
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = argtypes
    return fn


def call(name, *args):
    """Call a status-returning entry point; raise KrylovError on failure."""
    lib = load()
    st = getattr(lib, name)(*args)
    if st != CK_OK:
        msg = lib.ck_status_string(st).decode()
        if st == CK_ERR_UNSUPPORTED:
            raise UnsupportedOperator(name, st, msg)
        raise KrylovError(name, st, msg)


def q(name, *args):
    """Call a query entry point (sizes) and return its value."""
    return getattr(load(), name)(*args)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def version():
    return load().ck_version().decode()


def require_device(t, what="tensor"):
    """The engine runs on the HIP device only; CPU tensors are rejected loudly."""
    if not (isinstance(t, torch.Tensor) and t.is_cuda):
        raise ValueError(f"{what} must be a tensor on the HIP (cuda) device; the Krylov engine has no CPU path")
