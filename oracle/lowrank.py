"""NumPy restatement of the GP log-density caller around the SLQ path -- TEST INFRASTRUCTURE.

Follows, in the same operation order:

* ``low_rank.cholesky_partial`` (low_rank.py:56-99): column i of L is
  ``(A[:, i] - L L[i, :]) / sqrt(A[i, i] - <L[i], L[i]>)``;
* ``low_rank.cholesky_partial_pivot`` (low_rank.py:102-198): before step i the
  row with the largest ``|diag(A_p) - rowwise <L, L>|`` (first on ties, as
  ``argmax``) is swapped into position i (rows of L, the permutation), the
  step above runs on the permuted matrix, and the rows are finally put back
  in the original order (``L[argsort(P)]``); ``success`` records whether every
  ``l_ii^2`` was positive;
* ``low_rank.preconditioner`` (low_rank.py:7-53): the Woodbury solve
  ``(s I + L L^T)^{-1} v = v/s - U C^{-1} V v/s`` with ``U = L/sqrt(s)``,
  ``V = L^T/sqrt(s)`` and the capacitance ``C = I + V U`` factored by Cholesky;
* ``backend.linalg.cg`` (backend/linalg.py:107-108), i.e.
  ``jax.scipy.sparse.linalg.cg`` -- EXTERNAL: JAX is not installed and not
  vendored (``jax[cpu]`` unpinned, pyproject.toml:29).  Restated from JAX's
  published preconditioned CG (``_cg_solve``): ``x0 = 0``, ``r0 = b - A x0``,
  ``p0 = z0 = M r0``, ``gamma0 = r0 . z0``; stop when ``rs <= max(tol^2 b.b,
  atol^2)`` with ``rs = gamma`` without a preconditioner and ``r . r`` with one,
  or after ``maxiter = 10 n`` steps; per step ``alpha = gamma / (p . A p)``,
  ``x += alpha p``, ``r -= alpha A p``, ``z = M r``, ``gamma' = r . z``,
  ``p = z + (gamma' / gamma) p``.  Under ``vmap`` over right-hand sides a
  converged column keeps its state while the others continue;
* the log-density of tutorials/8_gaussian_logpdf.py:38-52:
  ``-(y . A^{-1} y + logdet A + n log(2 pi)) / 2`` with the SLQ estimate of
  ``logdet A`` and CG for ``A^{-1} y``.

All arrays are NumPy; nothing here touches a GPU.
"""

import numpy as np


def cholesky_partial(A, rank, dtype=np.float32):
    """``cholesky_partial(lambda i, j: A[i, j], nrows=n, rank=rank)()`` (low_rank.py:56-99)."""
    A = np.asarray(A, dtype=dtype)
    n = A.shape[0]
    if rank > n:
        raise ValueError(f"Rank exceeds n: {rank} >= {n}.")
    if rank < 1:
        raise ValueError(f"Rank must be positive, but {rank} < 1.")
    L = np.zeros((n, rank), dtype=dtype)
    for i in range(rank):
        l_ii = np.sqrt(A[i, i] - L[i] @ L[i])  # :91-92
        col = A[:, i] - L @ L[i, :]  # :94-95
        L[:, i] = col / l_ii  # :96-98
    return L, {}


def _swap(a, i, j):
    a[[i, j]] = a[[j, i]]


def cholesky_partial_pivot(A, rank, dtype=np.float32):
    """``cholesky_partial_pivot(lambda i, j: A[i, j], nrows=n, rank=rank)()`` (low_rank.py:102-198).

    Returns ``(L, {"success": bool}, P)``; ``P`` is the final permutation
    (position -> original row), which the reference keeps internal.
    """
    A = np.asarray(A, dtype=dtype)
    n = A.shape[0]
    if rank > n:
        raise ValueError(f"Rank exceeds nrows: {rank} >= {n}.")
    if rank < 1:
        raise ValueError(f"Rank must be positive, but {rank} < 1.")
    L = np.zeros((n, rank), dtype=dtype)
    P = np.arange(n)
    success = True
    for i in range(rank):
        diagonal = A[P, P]  # :153 (P_matrix and P receive the same swaps)
        residual = diagonal - np.einsum("ij,ij->i", L, L)  # :156
        k = int(np.argmax(np.abs(residual)))  # :157-158
        _swap(P, i, k)  # :161-163
        _swap(L, i, k)
        element = A[P[i], P[i]]  # :166
        column = A[P, P[i]]  # :167
        l_ii_squared = element - L[i] @ L[i]  # :173
        with np.errstate(invalid="ignore"):
            l_ii = np.sqrt(l_ii_squared)  # :174 (NaN when the matrix is exhausted)
        L[:, i] = (column - L @ L[i, :]) / l_ii  # :175-176, :180
        success = bool(success and l_ii_squared > 0.0)  # :177
    return L[np.argsort(P)], {"success": success}, P.copy()  # :127, :196-198


def preconditioner_solve(L, v, s, dtype=np.float32):
    """``preconditioner(cholesky)(v, s)`` for a factor ``L`` (n, rank) (low_rank.py:33-51)."""
    L = np.asarray(L, dtype=dtype)
    N, r = L.shape
    assert r <= N, (N, r)  # :38-39
    U = L / np.sqrt(dtype(s))  # :42
    V = L.T / np.sqrt(dtype(s))  # :43
    v = np.asarray(v, dtype=dtype) / dtype(s)  # :44
    C = np.eye(r, dtype=dtype) + V @ U  # :48-49
    R = np.linalg.cholesky(C.astype(np.float64))
    rhs = (V @ v).astype(np.float64)
    sol = np.linalg.solve(R.T, np.linalg.solve(R, rhs)).astype(dtype)  # :50
    return v - U @ sol  # :51


def cg(apply, b, tol=1e-5, atol=0.0, maxiter=None, M=None, dtype=np.float32):
    """``jax.scipy.sparse.linalg.cg(A, b, tol=..., atol=..., maxiter=..., M=...)[0]``
    for every column of ``b`` (P, n) as under ``vmap``.

    ``apply(X)`` maps a (P, n) block to ``A X`` row-wise; ``M`` likewise or
    None.  Returns ``(x (P, n), iterations (P,))``.
    """
    b = np.asarray(b, dtype=dtype)
    P, n = b.shape
    if maxiter is None:
        maxiter = 10 * n
    prec = (lambda r: r) if M is None else M
    bs = np.sum(b * b, axis=1)
    atol2 = np.maximum(np.square(dtype(tol)) * bs, np.square(dtype(atol)))
    x = np.zeros_like(b)
    r = b - apply(x)
    z = np.asarray(prec(r), dtype=dtype)
    p = z.copy()
    gamma = np.sum(r * z, axis=1)
    k = np.zeros(P, dtype=np.int64)

    def rs_of(r, gamma):
        return gamma if M is None else np.sum(r * r, axis=1)

    active = (rs_of(r, gamma) > atol2) & (k < maxiter)
    while active.any():
        Ap = np.asarray(apply(p), dtype=dtype)
        with np.errstate(invalid="ignore", divide="ignore"):  # frozen columns may divide 0 by 0
            alpha = gamma / np.sum(p * Ap, axis=1)
        x_ = x + alpha[:, None] * p
        r_ = r - alpha[:, None] * Ap
        z_ = np.asarray(prec(r_), dtype=dtype)
        gamma_ = np.sum(r_ * z_, axis=1)
        with np.errstate(invalid="ignore", divide="ignore"):
            beta_ = gamma_ / gamma
        p_ = z_ + beta_[:, None] * p
        a = active[:, None]
        x, r, p = np.where(a, x_, x), np.where(a, r_, r), np.where(a, p_, p)
        gamma = np.where(active, gamma_, gamma)
        k = k + active
        active = (rs_of(r, gamma) > atol2) & (k < maxiter)
    return x, k


def gaussian_logpdf(y, logdet, solve):
    """tutorials/8_gaussian_logpdf.py:46-51 with the log-determinant and solver given."""
    y = np.asarray(y, dtype=np.float64)
    quad = float(y @ np.asarray(solve(y), dtype=np.float64))
    return -0.5 * (quad + float(logdet) + y.shape[0] * np.log(2 * np.pi))
