"""Build libcolakrylov.so (gfx950) in-tree with hipcc.

Usage: python -m cola_amd.csrc.build [--jobs N] [--verbose]

Objects go to cola_amd/_build/, the shared library to cola_amd/lib/.  All HIP
sources are compiled with packed FP32 instructions disabled (hipcc's host pass
warns that it does not know the feature; that warning is harmless).
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIBNAME = "libcolakrylov.so"
ARCH = os.environ.get("COLA_AMD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-but-set-variable"]
DEVICE = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = ["ops.hip", "kron.hip", "cg.hip", "lanczos.hip", "tridiag.hip", "tridiag_bisect.hip", "trace.hip", "precond.hip", "mixed.hip",
           "learned.hip", "batched.hip", "fused_cg.hip", "gmres.hip", "symeig.hip", "coo.hip", "sketch.hip", "version.cpp"]


def lib_path():
    return os.path.join(LIBDIR, LIBNAME)


def _compile(src, verbose):
    path = os.path.join(HERE, src)
    obj = os.path.join(BUILD, src.replace(".", "_") + ".o")
    if src.endswith(".cpp"):
        cmd = ["g++", "-O2", "-fPIC", "-std=c++17", "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", path, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{res.stdout}\n{res.stderr}")
    warn = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if warn and verbose:
        print(warn, file=sys.stderr)
    return obj


def build(jobs=8, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    missing = [s for s in SOURCES if not os.path.exists(os.path.join(HERE, s))]
    if missing:
        raise FileNotFoundError(f"missing sources: {missing}")
    srcs = list(SOURCES)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    out = lib_path()
    tmp = out + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))


if __name__ == "__main__":
    main()
