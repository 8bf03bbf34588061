"""Writes tests/golden/full_config.json: the oracle's FP64 results on two full-size survey
workloads, for GPU parity tests that cannot run the oracle within a test's time:

* Lanczos: top-32 Ritz values of the 1024^2-grid Laplacian + 1e-3 I after 120 Lanczos
  steps (double Gram-Schmidt) from the seed-42 N(0,1) start vector;
* SLQ: the SLQ log-determinant of K1 x K2 x K3 + 0.01 I (64^3; each Ki an RBF kernel,
  lengthscale 0.1, on 64 sorted U[0,1] points from seed 0) with 32 Rademacher probes of the
  engine stream key 5 and 30 Lanczos steps.

Run: python tests/golden/make_full_config.py   (about two minutes on 8 host threads; no
reference code is imported or run -- the values come from the oracle restatement)."""
import json
import os
import sys

import numpy as np
import scipy.sparse as sp
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import lanczos as olz  # noqa: E402
from oracle import operators as oop  # noqa: E402
from oracle import rng as orng  # noqa: E402
from oracle import unary as oun  # noqa: E402


def kron64_factors(seed=0, ell=0.1):
    g = torch.Generator().manual_seed(seed)
    facs = []
    for _ in range(3):
        t = torch.sort(torch.rand(64, generator=g)).values
        facs.append(torch.exp(-(t[:, None] - t[None, :])**2 / (2 * ell**2)))
    return facs


def main():
    out = {}
    # Lanczos on the shifted Laplacian
    m = 1024
    ip, ix, v = oop.laplacian_2d_csr(m, 0.0)
    L = sp.csr_matrix((v.astype(np.float64), ix, ip), shape=(m * m, m * m))
    s0 = torch.randn(m * m, generator=torch.Generator().manual_seed(42))

    def mm(X):
        Xn = X.numpy()
        return torch.from_numpy(L @ Xn + 1e-3 * Xn)

    ev, _, _ = olz.lanczos_eigs(mm, m * m, s0.double(), 120, 1e-7, torch.float64)
    out["lanczos_laplacian_1024sq"] = {"start": "torch.randn(1024**2, generator=manual_seed(42))", "steps": 120,
                                       "tol": 1e-7, "top32_ritz_fp64": [float(x) for x in ev[-32:].tolist()]}
    # SLQ on the shifted Kronecker system
    facs = kron64_factors()
    n = 64**3
    Z = orng.rademacher(5, n, 32, torch.float64).T
    mk = lambda X: oop.kron_matmat([f.double() for f in facs], X) + 0.01 * X  # noqa: E731
    _, al, be, _ = olz.lanczos(mk, n, Z, 30, 1e-6, torch.float64)
    q, _, _ = oun.slq_quadratic_forms(al, be, (Z**2).sum(0), torch.log)
    out["slq_kron64"] = {"factors": "three RBF kernels, ell 0.1, on sorted torch.rand(64) of one seed-0 generator",
                         "shift": 0.01, "probes": 32, "key": 5, "steps": 30, "logdet_fp64": float(q.mean())}
    with open(os.path.join(HERE, "full_config.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out)[:400])


if __name__ == "__main__":
    main()
