"""The C ABI library loads and exports every function include/cola_krylov.h declares;
host-side size queries and status strings work without a GPU."""
import ctypes
import os
import re
import subprocess

import pytest

import cola_amd._native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cola_krylov.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\**\s+\**(ck_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(N.lib_path()):
        pytest.fail("libcolakrylov.so is not built (run python -m cola_amd.csrc.build)")
    return N.load()


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["ck_op_apply", "ck_cg_iterate", "ck_lanczos_iterate", "ck_arnoldi_post", "ck_tridiag_eigh_f64",
                 "ck_hutch_csr_bits", "ck_slq_quadrature", "ck_rng_rademacher_bits"]:
        assert must in names
    assert len(names) >= 39


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_every_bound_prototype_exists(lib):
    for name in N._PROTOS:
        assert hasattr(lib, name), name


def test_version_and_status(lib):
    assert N.version().startswith("cola_krylov")
    assert lib.ck_status_string(0) == b"ok"
    assert b"workspace" in lib.ck_status_string(4)
    assert lib.ck_status_string(99) == b"unknown status"


def test_size_queries(lib):
    assert N.q("ck_cg_stride", 1) == 64 and N.q("ck_cg_stride", 65) == 128
    assert N.q("ck_cg_state_bytes", 16, 1000) == 64 + 4 * (6 * 64 + 1002)
    assert N.q("ck_cg_nparts", 262144) == 64
    assert N.q("ck_cg_nparts", 1 << 30) == 1024
    assert N.q("ck_cg_partials_bytes", 262144, 16) == 3 * 4 * 16 * 64
    assert N.q("ck_lanczos_state_bytes", 1, 120) == 64 + 4 * 122
    assert N.q("ck_trace_workspace_bytes", 1000) == 8 * 4
    assert N.q("ck_dot_workspace_bytes", 10, 3) == 3 * 256 * 8


def test_op_workspace_queries(lib):
    op = N.CkOp()
    op.kind, op.n, op.nfactors = N.CK_OP_KRON_F32, 64**3, 3
    assert N.q("ck_op_workspace_bytes", op, 16) == 16 * 64**3 * 4
    op2 = N.CkOp()
    op2.kind, op2.n, op2.rank = N.CK_OP_WOODBURY_INV, 16384, 64
    # fp64 partials (one per 64-row F tile at rank 64: 256 tiles) and u = C^-1 F^T x
    assert N.q("ck_op_workspace_bytes", op2, 8) == 8 * (8 * 64 * 256 + 8 * 64)
    assert N.q("ck_mlp_param_count", 4, 64) == 4 * 128 + 128 + 128 * 128 + 128 + 128 * 64 + 64
    assert N.q("ck_mlp_param_count", 9, 64) == -1
    assert N.q("ck_mlp_grad_workspace_bytes", 16384, 4, 64) == 4 * 256 * 25408
    # one FP64 partial per 64-row chunk, at most 1024 partials (rows per partial grow past that)
    assert N.q("ck_gram_workspace_bytes", 1025, 64, 32) == 8 * 17 * 64 * 32
    assert N.q("ck_gram_workspace_bytes", 64 * 3072, 64, 32) == 8 * 1024 * 64 * 32


def test_bad_arguments_rejected_without_launch(lib):
    # null pointers / bad shapes return CK_ERR_BAD_ARG before any device work
    assert lib.ck_cg_start(0, 1, None, None, None, 1, None, ctypes.c_float(1e-6), 10, None, None) == 1
    assert lib.ck_tridiag_eigh_f64(0, 5, None, 5, None, 4, None, None, None, None, None) == 1
    assert lib.ck_tridiag_eigh_bisect_f64(1, 1025, None, 1025, None, 1024, None, None, None, None, None, 0, None) == 1
    assert N.q("ck_tridiag_eigh_bisect_workspace_bytes", 2, 120) == 8 * 2 * 4 * 120 * 128  # threads: m rounded to 64
    assert lib.ck_op_apply(None, 1, None, 1, None, 1, None, 0, None) == 1
    op = N.CkOp()
    op.kind, op.n = 99, 4
    buf = (ctypes.c_float * 4)()
    assert lib.ck_op_apply(ctypes.byref(op), 1, buf, 4, buf, 4, None, 0, None) == N.CK_ERR_UNSUPPORTED


@pytest.mark.parametrize("cname,pystruct", [("ck_op", N.CkOp), ("ck_bcg_plan", N.CkBcgPlan)])
def test_struct_layout_matches_header(tmp_path, cname, pystruct):
    """The ctypes struct layouts equal what the C compiler makes of include/cola_krylov.h."""
    fields = [f[0] for f in pystruct._fields_]
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {",
             f'  printf("%zu\\n", sizeof({cname}));']
    lines += [f'  printf("%zu\\n", offsetof({cname}, {f}));' for f in fields]
    lines += ["  return 0;", "}"]
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(pystruct)
    assert out[1:] == [getattr(pystruct, f).offset for f in fields]


def test_symmetric_flag_workspace():
    """ck_op.flags = CK_OP_SYMMETRIC adds the half-reading product's scratch to the operator
    workspace for orders that are a multiple of 128, and nothing otherwise."""
    N.load()

    def ws(n, flags, nrhs=8):
        op = N.CkOp()
        op.kind, op.n, op.scale, op.shift, op.flags = N.CK_OP_DENSE_F32, n, 1.0, 0.0, flags
        return N.q("ck_op_workspace_bytes", op, nrhs)

    assert ws(16384, N.CK_OP_SYMMETRIC) > ws(16384, 0) + 32 * 2**20  # row and column partial sums
    assert ws(1000, N.CK_OP_SYMMETRIC) == ws(1000, 0)  # not a multiple of 128: the full product
    assert ws(128, N.CK_OP_SYMMETRIC) > ws(128, 0)
    assert ws(16384, N.CK_OP_SYMMETRIC, 1) > 0
