"""CPU oracle for the Krylov hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain PyTorch (CPU) and NumPy, the algorithms of
wilson-labs/cola that the HIP engine in ``cola_amd`` replaces.  Every function
cites the reference file:line it follows.  It exists to *check* the GPU engine:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  Nothing in ``cola_amd`` imports or calls it, and the
product path never falls back to it.

How each part is pinned (what would catch a dropped term, wrong sign or index):

* ``rng``       -- SHA-256 key chain checked against the seed values quoted for
                   PRNGKey(42)/next_key/PRNGKey(0); the counter hash is checked
                   against independently computed SplitMix64 published values.
* ``operators`` -- CSR 3x4 known answer (tests/test_operators.py:170-204),
                   Kronecker vs ``torch.kron`` dense product, 1138_bus fixture vs
                   scipy CSR.
* ``cg``        -- diag(3,4,5) closed form, three-eigenvalue convergence in < 7
                   steps, the cond_fun truth table (tests/algorithms/test_cg.py),
                   FP64 dense solves.
* ``lanczos``   -- hand-computed cases case_early/case_3/case_2 with exact
                   alpha/beta/iteration counts (tests/algorithms/test_lanczos.py:268-300),
                   diag(4,2,1,6) max eigenvalue, A Q = Q T and Q^T Q = I invariants,
                   closed-form 2-D Laplacian spectrum.
* ``arnoldi``   -- A Q = Q H + h e_m^T invariant, Hessenberg structure, eigenvalues
                   of a triangular matrix.
* ``slq`` / ``unary`` -- exact log-determinant when the probes are scaled unit
                   vectors and Lanczos runs to full dimension; 10x10 known spectrum.
* ``hutch``     -- exact trace with Hadamard probes (sum z z^T = n I); closed-form
                   Laplacian trace.
* ``learned``   -- MLP gradients vs finite differences, energy-loss limits and
                   invariance, Adam vs torch.optim.Adam, pivoted Cholesky vs LAPACK
                   dpstrf, Nystrom exactness on low-rank K, Woodbury vs dense inverse.
* ``tridiag``   -- LAPACK (scipy.linalg.eigh_tridiagonal) in FP64.
* ``precond``   -- Woodbury identity against a dense FP64 inverse; pivoted Cholesky
                   against full Cholesky at full rank; MLP gradients against
                   torch autograd.
"""
