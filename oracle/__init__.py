"""CPU oracle for the SLQ log-determinant path -- TEST INFRASTRUCTURE ONLY.

This package is a NumPy restatement of matfree's stochastic Lanczos quadrature
path (Rademacher probes -> ``decomp.tridiag_sym(reortho="full")`` ->
``funm.monte_carlo_funm_sym_logdet`` -> ``stochtrace.estimator_monte_carlo``)
and of its reverse-mode adjoint.  Every function cites the reference file:line
it follows.

Who may use it: ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` -- always as the checker or as the timed
CPU baseline, never as part of the product.  ``matfree_amd`` never imports it.

Pinning (what keeps the restatement honest; tests live in
``tests/test_oracle_*.py``):

* ``prng``: Random123 threefry2x32-20 known-answer vectors and matfree's README
  known answer (10,000 Rademacher probes under key 1 give a trace estimate of
  exactly 504.0).  Bits of other keys/shapes follow from the same layout.
* ``slq.hessenberg_forward`` / ``slq.tridiag_sym``: Arnoldi relation,
  orthonormality and ``Q^T e0 = c v`` (the reference's own invariants), full
  rank reconstruction ``Q^T T Q = A``, Rayleigh-Ritz identity ``T = Q A Q^T``,
  and a single-step check with a non-orthonormal basis that fixes which CGS
  pass feeds ``H``.
* ``slq.quadform_logdet``: closed form ``x^T log(A) x`` at full depth and the
  Golub-Welsch identity; ``slq.estimator_mean``: dense ``slogdet``.
* ``slq.slq_logdet_vjp``: central finite differences of the FP64 forward and
  the dense closed form ``d/dtheta x^T log(A) x`` at full depth.
* ``slq.tridiag_eig_quadrature`` (the FP64 eigen/quadrature checker used for the
  GPU kernel): ``numpy.linalg.eigh`` of the dense tridiagonal matrix.
* ``slq.lanczos_forward`` / ``slq.lanczos_adjoint`` (``reortho="none"``): the
  reference's decomposition relation and reconstruction, agreement with
  Arnoldi in FP64 while orthogonal, central differences for random
  cotangents on every output, planted mistakes.
* ``lowrank`` (partial Cholesky with and without pivoting, the Woodbury
  preconditioner, CG with JAX's stopping rule, the Gaussian log-density of the
  GP tutorial): the reference's low-rank tests, ``numpy.linalg.cholesky``,
  dense solves, CG's termination in at most d steps for d distinct
  eigenvalues, ``scipy.stats.multivariate_normal.logpdf``, planted mistakes.

No function is left "parity unpinned" except the bitwise XLA summation order
and LAPACK ``ssyevd`` last bits, which only an import of JAX could fix (JAX is
not installed here; see the module docstrings).
"""
