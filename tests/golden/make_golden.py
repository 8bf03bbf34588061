"""Writes tests/golden/kat.json: the known-answer values the reference's own tests
state by hand (cited per entry) plus closed-form values used to pin the oracle.

Run: python tests/golden/make_golden.py   (no reference code is imported or run)."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def laplacian_eigs(m, shift):
    j = np.arange(1, m + 1)
    lam = 2.0 - 2.0 * np.cos(np.pi * j / (m + 1))
    return np.sort((lam[:, None] + lam[None, :]).ravel() + shift)


def main():
    kat = {
        "cg_diag345": {"source": "tests/algorithms/test_cg.py:170-190", "A_diag": [3., 4., 5.], "b": [1., 1., 1.],
                       "x": [1 / 3, 1 / 4, 1 / 5]},
        "cg_cond_fun": {"source": "tests/algorithms/test_cg.py:107-124", "cases": [
            {"res": [[1.0, 0.1, 0.0001], [1.0, 0.1, 0.0001]], "k": 5, "tol": 1e-2, "max_iters": 10, "flag": True},
            {"res": [[1.0, 1.0, 1.0], [1.0, 1.0, 1.0]], "k": 11, "tol": 1e-2, "max_iters": 10, "flag": False},
            {"res": [[1.0, 1.0, 1.0], [1.0, 1.0, 1.0]], "k": 3, "tol": 10.0, "max_iters": 10, "flag": False}]},
        "cg_repeated_eig": {"source": "tests/algorithms/test_cg.py:148-167",
                            "diag": [1.0] * 10 + [0.5] * 10 + [0.25] * 10, "max_k_exclusive": 7},
        "lanczos_cases": {"source": "tests/algorithms/test_lanczos.py:268-300", "cases": [
            {"name": "early", "A": [[4, 0, 0], [0, 2, 0], [0, 0, 1]], "rhs": [1, 0, 0], "beta": [4], "alpha": [],
             "iters": 1},
            {"name": "3", "A": [[7, 0, 1.0], [0, 1, 0.1], [1.0, 0.1, 3]], "rhs": [0, 1, 0], "beta": [1, 3, 7],
             "alpha": [0.1, 1.0], "iters": 3},
            {"name": "2", "A": [[1, 0.1, 0], [0.1, 2, 0.1], [0, 0.1, 4]], "rhs": [1, 0, 0], "beta": [1, 2, 4],
             "alpha": [0.1, 0.1], "iters": 3}]},
        "lanczos_max_eig": {"source": "tests/algorithms/test_lanczos.py:195-217", "diag": [4, 2, 1, 6], "max": 6},
        "csr_3x4": {"source": "tests/test_operators.py:170-204", "data": [4., 5., 6., 1., 2., 3.],
                    "rows": [2, 2, 2, 0, 0, 1], "cols": [0, 1, 2, 1, 3, 3], "shape": [3, 4],
                    "x1": [1., 0., -1., 1.], "Ax1": [2., 3., -2.], "x2": [1., 2., 3.], "x2A": [12., 16., 18., 8.]},
        "kron_2x3_4x2": {"source": "tests/test_operators.py:527-541", "A": [[1., 2., 7], [3., 4., 5]],
                         "B": [[-1, 4.], [3., 3.], [2., -1.], [2., 1.]]},
        "prng_keys": {"source": "cola/backends/torch_fns.py:154-155,222-230 (values quoted with the survey)",
                      "PRNGKey42": 3300923923, "next_key_of_42": 1419105812, "PRNGKey0": 1514845414},
        "splitmix64_seed0": {"source": "SplitMix64 reference outputs (Steele, Lea, Flood 2014), state 0",
                             "outputs_hex": ["e220a8397b1dcdaf", "6e789e6aa1b965f4", "06c45d188009454f"]},
        "laplacian_1024": {"source": "closed form, 5-point Dirichlet Laplacian on 1024x1024, shift 1e-3",
                           "top32": laplacian_eigs(1024, 1e-3)[-32:].tolist(),
                           "min": float(laplacian_eigs(1024, 1e-3)[0]), "trace": 4.001 * 1024 * 1024,
                           "nnz": 5 * 1024 * 1024 - 4 * 1024},
    }
    with open(os.path.join(HERE, "kat.json"), "w") as f:
        json.dump(kat, f, indent=1)


if __name__ == "__main__":
    main()
