"""The Gaussian log-density of tutorials/8_gaussian_logpdf.py:38-52 on the GPU.

``create_logpdf(matvec, vec_like, num_matvecs=6, num_samples=500)`` bundles
the SLQ log-determinant estimator and CG for the linear system into
``logpdf(y, key, *params) = -(y . A^{-1} y + logdet A + n log(2 pi)) / 2``.
It is differentiable in the operator parameters through torch autograd (the
counterpart of ``jax.grad`` in the tutorial): the log-determinant through the
SLQ adjoint, the quadratic form through ``-x^T (dA) x`` at ``x = A^{-1} y``
(JAX differentiates its CG implicitly with a second CG solve of the same
system; the converged first solve is reused here).
"""

import math

import torch

from . import _runtime, decomp, funm, linalg, stochtrace


class _QuadForm(torch.autograd.Function):
    """``y^T A(params)^{-1} y`` by CG, with its parameter and ``y`` cotangents."""

    @staticmethod
    def forward(ctx, matvec, y, tol, maxiter, *params):
        x, _ = linalg.cg_batched(matvec, y, params, tol=tol, maxiter=maxiter)
        yy = _runtime.as_f32(y, x.device)
        quad = torch.dot(yy.to(torch.float64), x.to(torch.float64))
        ctx.matvec = matvec
        ctx.params = params
        ctx.save_for_backward(x)
        return quad

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        params = ctx.params
        grads = []
        if any(isinstance(p, torch.Tensor) and p.requires_grad for p in params):
            cot = linalg.quadform_vjp(ctx.matvec, x, params)
            for p, c in zip(params, cot):
                if isinstance(p, torch.Tensor):
                    grads.append((-g * c.to(torch.float64)).to(p.dtype).reshape(p.shape))
                else:
                    grads.append(None)
        else:
            grads = [None] * len(params)
        gy = None
        if ctx.needs_input_grad[1]:
            gy = 2.0 * g * x.to(torch.float64)
        return (None, gy, None, None, *grads)


def quadform_inverse(matvec, y, *params, tol=1e-5, maxiter=None):
    """``y^T A(params)^{-1} y`` (float64 tensor), differentiable in ``params`` and ``y``."""
    return _QuadForm.apply(matvec, y, tol, maxiter, *params)


def create_logpdf(matvec, vec_like, *, num_matvecs=6, num_samples=500, tol=1e-5):
    """Create a matrix-free Gaussian logpdf (tutorials/8_gaussian_logpdf.py:38-52)."""
    tridiag = decomp.tridiag_sym(num_matvecs)
    integrand = funm.monte_carlo_funm_sym_logdet(tridiag)
    sampler = stochtrace.sampler_signs(vec_like, num=num_samples)
    estimator = stochtrace.estimator_monte_carlo(integrand, sampler=sampler)

    def logpdf(y, key, *params):
        dim = len(y)
        logdet = estimator(matvec, key, *params)
        quad = quadform_inverse(matvec, y, *params, tol=tol)
        return -0.5 * (quad + logdet + dim * math.log(2 * math.pi))

    return logpdf
