"""Benchmark of the Krylov engine on one MI355X.

Default workload (the CG-solves/sec metric): batched CG on the Kronecker SPD
operator K1 x K2 x K3 + 0.01 I, N = 64^3 = 262144, 16 N(0,1) right-hand sides,
each Ki a 64x64 RBF kernel (lengthscale 0.1) on 64 sorted Uniform[0,1] points.
A step is one full batched solve to the target relative residual, checked in
FP64 for every right-hand side; value = solved right-hand sides per second.

    python bench.py [--steps K] [--warmup W] [--workload kron_cg|dense_cg|lanczos|hutch|slq]

Prints one JSON line with the metric, a roofline object for the dominant kernel
(measured with HIP events on the engine's stream) and a CPU baseline (the CPU
torch restatement of the reference algorithm, timed on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import cola_amd as ca  # noqa: E402
import cola_amd._native as N  # noqa: E402
from cola_amd.linalg import run_batched_cg  # noqa: E402

METRIC = "CG solves/sec to 1e-6 rel. residual: learned+mixed-prec vs plain; HBM GB/s"
HBM_PEAK = 8000.0  # GB/s
FP32_MFMA_PEAK = 157.3  # TFLOP/s


def rbf_factor(t, ell):
    return torch.exp(-(t[:, None] - t[None, :])**2 / (2 * ell**2))


def kron_workload(dev, n1=64, nrhs=16, seed=0):
    g = torch.Generator().manual_seed(seed)
    facs = [rbf_factor(torch.sort(torch.rand(n1, generator=g)).values, 0.1) for _ in range(3)]
    b = torch.randn(n1**3, nrhs, generator=g)
    K = ca.Kronecker(*[ca.Dense(f.to(dev)) for f in facs])
    A = ca.PSD(K + 0.01 * ca.I_like(K))
    return A, facs, b


def kron_apply_f64(facs, X):
    """FP64 (K1 x K2 x K3 + 0.01 I) X on the device, for the residual check only."""
    dims = [f.shape[0] for f in facs]
    ev = X.double().reshape(*dims, -1)
    for i, f in enumerate(facs):
        M = f.double().to(X.device)
        front = torch.movedim(ev, i, 0)
        shape = front.shape
        ev = torch.movedim((M @ front.reshape(shape[0], -1)).reshape(shape), 0, i)
    return ev.reshape(X.shape[0], -1) + 0.01 * X.double()


def time_kernel(fn, reps=50):
    """Average duration of fn's launches (HIP events on the current stream)."""
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps * 1e-3  # seconds


def cpu_threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count()


def bench_kron_cg(args, dev):
    A, facs, b = kron_workload(dev)
    n, B = b.shape
    bd = b.to(dev)
    tol = args.tol
    # warmup (includes first-touch allocation)
    for _ in range(args.warmup):
        x, r, k, info = run_batched_cg(A, bd, None, 1000, tol, None, chunk=args.chunk)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        x, r, k, info = run_batched_cg(A, bd, None, 1000, tol, None, chunk=args.chunk)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    # FP64 check of every right-hand side
    res = (kron_apply_f64(facs, x) - bd.double()).norm(dim=0) / bd.double().norm(dim=0)
    max_res = float(res.max())

    # roofline of the Kronecker matvec kernels, timed alone on the solve's block shape
    op, keep = A.hip_op()
    ld = (n + 63) // 64 * 64
    Xb = torch.randn(B, ld, device=dev)
    Yb = torch.empty_like(Xb)
    wb = N.q("ck_op_workspace_bytes", op, B)
    work = torch.empty(wb // 4, device=dev)
    s = N.stream_ptr(dev)
    t_mv = time_kernel(lambda: N.call("ck_op_apply", op, B, N.ptr(Xb), ld, N.ptr(Yb), ld, N.ptr(work), wb, s))
    flops_mv = 3 * 2 * 64 * n * B  # three mode products
    bytes_mv = 4 * n * B * 5  # pass A: read X, write U; pass B: read U, read X, write Y
    # CG vector update kernel (x += a p, r -= a Ap with partial sums), same block shape
    from cola_amd.linalg.cg import CGState
    st = CGState(n, B, 10, dev, False)
    st.p.normal_()
    st.Ap.normal_()
    st.r.normal_()
    N.call("ck_cg_start", n, B, N.ptr(st.r), N.ptr(st.r), N.ptr(st.p), ld, N.ptr(st.state), 1e-30, 10**9,
           N.ptr(st.partials), s)
    t_xr = time_kernel(lambda: N.call("ck_cg_step_xr", n, B, N.ptr(st.p), N.ptr(st.Ap), N.ptr(st.x), N.ptr(st.r), ld,
                                      N.ptr(st.state), N.ptr(st.partials), s))
    bytes_xr = 4 * n * B * 7  # partial p.Ap (2 reads) + update (4 reads, 2 writes) -- two launches
    roof = {"bound": "mfma", "kernel": "kron matvec (kron_pass_a + kron_pass_b)",
            "achieved": flops_mv / t_mv / 1e12, "peak": FP32_MFMA_PEAK, "unit": "TFLOP/s",
            "frac": flops_mv / t_mv / 1e12 / FP32_MFMA_PEAK, "traffic": None,
            "launch_us": t_mv * 1e6, "algorithmic_flops_per_launch": flops_mv,
            "algorithmic_bytes_per_launch": bytes_mv, "hbm_gbs_equiv": bytes_mv / t_mv / 1e9,
            "cg_update_xr_us": t_xr * 1e6, "cg_update_xr_gbs": bytes_xr / t_xr / 1e9}

    out = {"metric": METRIC, "value": B / dt, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded RBF Kronecker factors, N(0,1) rhs)",
           "config": {"workload": "kron_cg: K1xK2xK3+0.01I, N=64^3=262144, 16 rhs, plain FP32 CG",
                      "global_batch": B, "seq_len": n, "parallelism": "single GPU", "tol": tol,
                      "cg_iterations": int(k), "max_fp64_rel_residual": max_res},
           "roofline": roof}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_kron(facs, b, k, tol)
    return out


def cpu_baseline_kron(facs, b, k_gpu, tol, iters=20):
    """The CPU torch restatement of cola's run_batched_cg with Kronecker._matmat, timed
    for `iters` CG steps on the full 16-rhs system and scaled to the k_gpu steps a
    solve takes."""
    from oracle import cg as ocg
    from oracle import operators as oop
    torch.set_num_threads(cpu_threads())
    mm = oop.shifted(lambda V: oop.kron_matmat(facs, V), 0.01)
    ocg.run_batched_cg(mm, b, None, 2, tol, None)  # warm-up
    t0 = time.perf_counter()
    _, _, kk, _ = ocg.run_batched_cg(mm, b, None, iters, tol, None)
    dt = (time.perf_counter() - t0) / max(kk, 1)
    solve_s = dt * k_gpu
    return {"value": b.shape[1] / solve_s, "unit": "solves/s", "cores": cpu_threads(), "kind": "port",
            "sample": f"{kk} CG steps of the 16-rhs 64^3 Kronecker solve ({dt * 1e3:.1f} ms/step, fp32 torch), "
                      f"scaled to the {int(k_gpu)} steps of a full solve"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="kron_cg")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--chunk", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus > 1 is not supported"}))
        sys.exit(2)
    dev = torch.device("cuda:0")
    N.load()
    fn = {"kron_cg": bench_kron_cg}[args.workload]
    out = fn(args, dev)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
