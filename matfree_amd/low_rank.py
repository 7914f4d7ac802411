"""Partial Cholesky factorisations and the low-rank preconditioner (low_rank.py:7-198).

``cholesky_partial`` / ``cholesky_partial_pivot`` keep the reference's
factory signatures and error messages; the element function ``mat_el`` must
be one of the package's operators, read as the element function of its
matrix (``operators.dense_matvec``: ``mat_el(i, j, A) = A[i, j]``;
``operators.RBF(points)``: ``mat_el(i, j, lengthscale, noise) =
exp(-|x_i - x_j|^2 / (2 l^2)) + noise [i == j]``).  A plain Python callable
is rejected with ``TypeError``, as for matvecs.  The factorisation runs in
libslq's HIP kernels (``slq_cholesky_partial``); ``preconditioner`` applies
``(s I + L L^T)^{-1}`` by the Woodbury identity (``slq_lowrank_solve``).
"""

import ctypes

import torch

from . import _runtime, operators


class _Cholesky:
    def __init__(self, mat_el, nrows, rank, pivot):
        self.mat_el = mat_el
        self.nrows = int(nrows)
        self.rank = int(rank)
        self.pivot = pivot

    def _check(self):
        if self.rank > self.nrows:
            word = "nrows" if self.pivot else "n"
            raise ValueError(f"Rank exceeds {word}: {self.rank} >= {self.nrows}.")
        if self.rank < 1:
            raise ValueError(f"Rank must be positive, but {self.rank} < {1}.")

    def factor(self, *params):
        """``(L (nrows, rank), success (bool tensor), perm (nrows,) int64)``."""
        self._check()
        op = operators.resolve(self.mat_el)
        device = _runtime.target_device(*params)
        desc, keep = op.descriptor(self.nrows, params, device)
        L = torch.empty((self.nrows, self.rank), dtype=torch.float32, device=device)
        perm = torch.empty((self.nrows,), dtype=torch.int64, device=device)
        success = torch.empty((), dtype=torch.int32, device=device)
        ctx = _runtime.context(device)
        ctx.call("slq_cholesky_partial", ctypes.byref(desc), self.rank, 1 if self.pivot else 0,
                 _runtime.ptr(L), _runtime.ptr(perm), _runtime.ptr(success))
        del keep
        return L, success.bool(), perm

    def __call__(self, *params):
        L, success, _ = self.factor(*params)
        return L, ({"success": success} if self.pivot else {})


def cholesky_partial(mat_el, /, *, nrows, rank):
    """Partial Cholesky factorisation without pivoting (low_rank.py:56-99)."""
    return _Cholesky(mat_el, nrows, rank, pivot=False)


def cholesky_partial_pivot(mat_el, /, *, nrows, rank):
    """Partial Cholesky factorisation with pivoting (low_rank.py:102-198)."""
    return _Cholesky(mat_el, nrows, rank, pivot=True)


def lowrank_solve(L, V, s):
    """``(s I + L L^T)^{-1}`` applied to ``V`` ((n,) or (P, n)) for ``L`` (n, rank)."""
    device = _runtime.target_device(L, V)
    L = _runtime.as_f32(L, device)
    n, rank = L.shape
    if rank > n:
        raise AssertionError((n, rank))  # low_rank.py:38-39
    x = _runtime.as_f32(V, device)
    single = x.ndim == 1
    X = x.reshape(1, -1) if single else x
    out = torch.empty_like(X)
    ctx = _runtime.context(device)
    ctx.call("slq_lowrank_solve", _runtime.ptr(L), n, rank, float(_runtime.scalar(s)),
             _runtime.ptr(X), X.shape[0], _runtime.ptr(out))
    return out[0] if single else out


class LowRankShift:
    """``(s I + L L^T)^{-1}`` as the ``M`` of :func:`matfree_amd.linalg.cg`."""

    def __init__(self, L, s):
        self.L = L
        self.s = float(_runtime.scalar(s))

    def __call__(self, v):
        return lowrank_solve(self.L, v, self.s)


class _Preconditioner:
    def __init__(self, cholesky):
        self.cholesky = cholesky

    def __call__(self, v, s, *cholesky_params):
        chol, info = self.cholesky(*cholesky_params)
        return lowrank_solve(chol, v, s), info

    def bind(self, s, *cholesky_params):
        """Factor once and return the solve as a CG preconditioner (``M``)."""
        chol, _ = self.cholesky(*cholesky_params)
        return LowRankShift(chol, s)


def preconditioner(cholesky, /):
    """Turn a low-rank approximation into a preconditioner (low_rank.py:7-53).

    Returns ``solve(v, s, *params) -> ((s I + L L^T)^{-1} v, info)`` with
    ``L`` from ``cholesky(*params)``; ``solve.bind(s, *params)`` gives the
    same operator, factored once, for :func:`matfree_amd.linalg.cg`.
    """
    return _Preconditioner(cholesky)
