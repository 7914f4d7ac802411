"""matfree_amd: matfree's SLQ log-determinant path on AMD Instinct MI355X (gfx950).

Drop-in for the reference API on the path
``stochtrace.sampler_signs -> decomp.tridiag_sym(reortho="full") ->
funm.monte_carlo_funm_sym_logdet -> stochtrace.estimator_monte_carlo``;
the arithmetic runs in the hand-written HIP kernels of ``libslq.so``
(C ABI in ``include/slq.h``).
"""

from . import decomp, funm, gp, linalg, low_rank, operators, prng, serving, stochtrace

__all__ = ["decomp", "funm", "gp", "linalg", "low_rank", "operators", "prng", "serving",
           "stochtrace"]
