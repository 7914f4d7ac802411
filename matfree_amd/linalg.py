"""Conjugate gradients: ``backend.linalg.cg`` (backend/linalg.py:107-108).

matfree calls ``jax.scipy.sparse.linalg.cg(Av, b)``; here ``cg(matvec, b,
*params)`` runs the same iteration (x0 = 0, stopping rule
``|r|^2 <= max(tol^2 |b|^2, atol^2)`` -- ``r . z`` in place of ``|r|^2`` without
a preconditioner, as JAX does -- and ``maxiter = 10 n`` by default) in libslq's
HIP kernels (``slq_cg``) with the package's block matvec.  A batch of
right-hand sides (P, n) is solved at once with JAX's ``vmap`` semantics: a
converged column keeps its state while the others iterate.
"""

import ctypes

import torch

from . import _runtime, low_rank, operators


def cg_batched(matvec, B, params, *, tol=1e-5, atol=0.0, maxiter=None, M=None):
    """Solve ``A(params) X = B`` row by row; returns ``(X (P, n), iterations (P,))``."""
    op = operators.resolve(matvec)
    device = _runtime.target_device(B, *params)
    Bt = _runtime.as_f32(B, device)
    single = Bt.ndim == 1
    Bt = Bt.reshape(1, -1) if single else Bt
    P, n = Bt.shape
    if maxiter is None:
        maxiter = 10 * n
    L, rank, s = None, 0, 1.0
    if M is not None:
        if not isinstance(M, low_rank.LowRankShift):
            raise TypeError("M must come from low_rank.preconditioner(...).bind(s, *params)")
        L = _runtime.as_f32(M.L, device)
        rank, s = int(L.shape[1]), M.s
    desc, keep = op.descriptor(n, params, device)
    X = torch.empty_like(Bt)
    iters = torch.empty((P,), dtype=torch.int64, device=device)
    ctx = _runtime.context(device)
    ctx.call("slq_cg", ctypes.byref(desc), _runtime.ptr(Bt), P, float(tol), float(atol),
             int(maxiter), _runtime.ptr(L), rank, float(s), _runtime.ptr(X), _runtime.ptr(iters))
    del keep
    return (X[0], iters[0]) if single else (X, iters)


def cg(matvec, b, *params, tol=1e-5, atol=0.0, maxiter=None, M=None):
    """``jax.scipy.sparse.linalg.cg`` for ``A(params) x = b``: returns ``(x, None)``."""
    x, _ = cg_batched(matvec, b, params, tol=tol, atol=atol, maxiter=maxiter, M=M)
    return x, None


def quadform_vjp(matvec, X, params):
    """Cotangents of ``sum_p x_p^T A(params) x_p`` w.r.t. the operator parameters."""
    op = operators.resolve(matvec)
    device = _runtime.target_device(X, *params)
    Xt = _runtime.as_f32(X, device)
    Xt = Xt.reshape(1, -1) if Xt.ndim == 1 else Xt
    P, n = Xt.shape
    desc, keep = op.descriptor(n, params, device)
    ctx = _runtime.context(device)
    if isinstance(op, operators.RBF):
        d = torch.empty((2,), dtype=torch.float64, device=device)
        ctx.call("slq_quadform_vjp", ctypes.byref(desc), _runtime.ptr(Xt), P, _runtime.ptr(d), None)
        out = (d[0], d[1])
    else:
        D = torch.empty((n, n), dtype=torch.float32, device=device)
        ctx.call("slq_quadform_vjp", ctypes.byref(desc), _runtime.ptr(Xt), P, None, _runtime.ptr(D))
        out = (D,)
    del keep
    return out
