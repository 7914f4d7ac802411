"""SLQ integrands on the GPU (matfree.funm subset on the log-determinant path).

* :func:`monte_carlo_funm_sym_logdet` (funm.py:186-202), also exported as
  :data:`integrand_funm_sym_logdet`;
* :func:`monte_carlo_funm_sym` (funm.py:205-243) and
  :func:`dense_funm_sym_eigh` (funm.py:322-335) for ``matfun = log``;
* :func:`funm_lanczos_sym` (funm.py:114-147): ``f(A) v`` for log, exp, sin,
  cos, sqrt and reciprocal.

The integrand ``quadform(matvec, v0, *params) = |v0|^2 [f(T)]_00`` runs as one
libslq call (``slq_logdet``): normalise, Arnoldi with full
reorthogonalisation, FP64 tridiagonal eigen-solve and Gauss quadrature.
"""

import ctypes
import math

import numpy as np
import torch

from . import _lib, _runtime, _tree, decomp, operators

# The scalar functions the FP64 tridiagonal kernels apply to Ritz values
# (NumPy, math and torch spellings of each).
_MATFUN_CODES = [
    ((np.log, math.log, torch.log), _lib.MATFUNS["log"]),
    ((np.exp, math.exp, torch.exp), _lib.MATFUNS["exp"]),
    ((np.sin, math.sin, torch.sin), _lib.MATFUNS["sin"]),
    ((np.cos, math.cos, torch.cos), _lib.MATFUNS["cos"]),
    ((np.sqrt, math.sqrt, torch.sqrt), _lib.MATFUNS["sqrt"]),
    ((np.reciprocal, torch.reciprocal), _lib.MATFUNS["reciprocal"]),
]


class _DenseFunmSymEigh:
    """``dense_funm_sym_eigh(matfun)`` (funm.py:322-335)."""

    def __init__(self, matfun):
        self.matfun = matfun

    @property
    def code(self):
        """The libslq code of ``matfun``, or None for a function it does not apply."""
        for fns, code in _MATFUN_CODES:
            if any(self.matfun is f for f in fns):
                return code
        return None

    @property
    def is_log(self):
        return self.code == _lib.MATFUNS["log"]

    def __call__(self, dense_matrix):
        raise NotImplementedError(
            "matfree_amd fuses f(T) into its FP64 tridiagonal eigen-solve kernel; "
            "use it through monte_carlo_funm_sym / monte_carlo_funm_sym_logdet"
        )


def dense_funm_sym_eigh(matfun):
    """Matrix function via symmetric eigendecomposition (funm.py:322-335)."""
    return _DenseFunmSymEigh(matfun)


def logdet_per_probe(matvec, Z, num_matvecs, params, want_coeffs=False, reortho="full"):
    """Run ``slq_logdet`` (``reortho="full"``) or ``slq_logdet_lanczos``
    (``reortho="none"``, plain Lanczos) on probes ``Z`` (P, n).

    Returns ``(per_probe (P,) float64, mean () float64)`` and, with
    ``want_coeffs``, the tridiagonal ``(alpha (P, k), beta (P, k-1))``.
    """
    op = operators.resolve(matvec)
    shape = tuple(Z.shape)
    P, n = (1, shape[0]) if len(shape) == 1 else shape
    k = int(num_matvecs)
    if k < 0 or k > n:
        raise ValueError(decomp._error_num_matvecs(k, maxval=n, minval=0))
    device = _runtime.target_device(Z, *params)
    Z = _runtime.as_f32(Z, device).reshape(P, n)
    desc, keep = op.descriptor(n, params, device)
    per_probe = torch.empty((P,), dtype=torch.float64, device=device)
    mean = torch.empty((), dtype=torch.float64, device=device)
    alpha = beta = None
    if want_coeffs:
        alpha = torch.empty((P, k), dtype=torch.float32, device=device)
        beta = torch.empty((P, max(k - 1, 0)), dtype=torch.float32, device=device)
    ctx = _runtime.context(device)
    ctx.call(
        "slq_logdet" if reortho == "full" else "slq_logdet_lanczos",
        ctypes.byref(desc),
        _runtime.ptr(Z),
        P,
        k,
        _runtime.ptr(per_probe),
        _runtime.ptr(mean),
        _runtime.ptr(alpha) if want_coeffs else None,
        _runtime.ptr(beta) if want_coeffs and k > 1 else None,
    )
    del keep
    if want_coeffs:
        return per_probe, mean, alpha, beta
    return per_probe, mean


def logdet_value_and_grad(matvec, Z, num_matvecs, params, reortho="full"):
    """Run ``slq_logdet_vjp`` (``reortho="full"``: the Arnoldi adjoint) or
    ``slq_logdet_lanczos_vjp`` (``reortho="none"``: the Lanczos adjoint) on
    probes ``Z`` (P, n).

    Returns ``(per_probe (P,), mean (), grads)`` where ``grads`` holds one
    cotangent of the probe mean per parameter in ``params``: for
    :class:`operators.RBF` the float64 scalars ``d mean / d lengthscale`` and
    ``d mean / d noise``.  This is reverse mode through the estimator
    (decomp.py:398-422, 480-600 or 181-197, 268-348; funm.py:331-333;
    stochtrace.py:50).
    """
    op = operators.resolve(matvec)
    shape = tuple(Z.shape)
    P, n = (1, shape[0]) if len(shape) == 1 else shape
    k = int(num_matvecs)
    device = _runtime.target_device(Z, *params)
    Z = _runtime.as_f32(Z, device).reshape(P, n)
    desc, keep = op.descriptor(n, params, device)
    per_probe = torch.empty((P,), dtype=torch.float64, device=device)
    mean = torch.empty((), dtype=torch.float64, device=device)
    if isinstance(op, operators.RBF):
        dparams = torch.empty((2,), dtype=torch.float64, device=device)
        dmatrix = None
    else:
        dparams = None
        dmatrix = torch.empty((n, n), dtype=torch.float32, device=device)
    ctx = _runtime.context(device)
    ctx.call(
        "slq_logdet_vjp" if reortho == "full" else "slq_logdet_lanczos_vjp",
        ctypes.byref(desc),
        _runtime.ptr(Z),
        P,
        k,
        _runtime.ptr(per_probe),
        _runtime.ptr(mean),
        _runtime.ptr(dparams),
        _runtime.ptr(dmatrix),
    )
    del keep
    grads = (dparams[0], dparams[1]) if dparams is not None else (dmatrix,)
    return per_probe, mean, grads


def _needs_grad(params):
    return torch.is_grad_enabled() and any(
        isinstance(p, torch.Tensor) and p.requires_grad for p in params
    )


class _LogdetMean(torch.autograd.Function):
    """The probe mean as a differentiable function of the operator parameters.

    The forward pass runs ``slq_logdet_vjp`` (value and cotangents in one
    call, as a custom VJP stores its residuals); the backward pass scales the
    stored cotangents by the incoming gradient.
    """

    @staticmethod
    def forward(ctx, matvec, samples, num_matvecs, reortho, *params):
        _, mean, grads = logdet_value_and_grad(matvec, samples, num_matvecs, params, reortho)
        shaped = []
        for p, g in zip(params, grads):
            if isinstance(p, torch.Tensor):
                g = g.to(dtype=p.dtype, device=p.device).reshape(p.shape)
            shaped.append(g)
        ctx.save_for_backward(*shaped)
        return mean

    @staticmethod
    def backward(ctx, gmean):
        grads = ctx.saved_tensors
        return (None, None, None, None, *[(gmean.to(g.dtype) * g) for g in grads])


class _QuadFormSym:
    """``monte_carlo_funm_sym(...).quadform`` (funm.py:226-241)."""

    def __init__(self, dense_funm, tridiag_sym):
        if not isinstance(dense_funm, _DenseFunmSymEigh) or not dense_funm.is_log:
            raise NotImplementedError(
                "matfree_amd's GPU integrand implements dense_funm_sym_eigh(log) only"
            )
        self.reortho = _reortho_of(tridiag_sym)
        self.tridiag_sym = tridiag_sym
        self.num_matvecs = tridiag_sym.num_matvecs

    def batched(self, matvec, samples, *parameters):
        """Per-probe values ``|z_p|^2 [log T_p]_00`` for the rows of ``samples``."""
        per_probe, _ = logdet_per_probe(matvec, samples, self.num_matvecs, parameters,
                                        reortho=self.reortho)
        return per_probe

    def estimate(self, matvec, samples, *parameters):
        """Mean over probes, reduced on the device.

        Differentiable in the operator parameters: with any parameter a
        tensor that requires grad, the result carries a backward that uses
        the GPU adjoint of the decomposition (Arnoldi for ``reortho="full"``,
        Lanczos for ``"none"``; torch's counterpart of ``jax.grad`` through
        the estimator).
        """
        if _needs_grad(parameters):
            return _LogdetMean.apply(matvec, samples, self.num_matvecs, self.reortho, *parameters)
        _, mean = logdet_per_probe(matvec, samples, self.num_matvecs, parameters,
                                   reortho=self.reortho)
        return mean

    def __call__(self, matvec, v0, *parameters):
        flat, _ = _tree.ravel(v0, decomp._host_or_gpu(v0, *parameters))
        if _needs_grad(parameters):
            return _LogdetMean.apply(matvec, flat[None, :], self.num_matvecs, self.reortho,
                                     *parameters)
        per_probe, _ = logdet_per_probe(matvec, flat[None, :], self.num_matvecs, parameters,
                                        reortho=self.reortho)
        return per_probe[0]


def _reortho_of(tridiag_sym):
    if isinstance(tridiag_sym, decomp._TridiagSymFull):
        return "full"
    if isinstance(tridiag_sym, decomp._TridiagSymNone):
        return "none"
    raise TypeError("pass decomp.tridiag_sym(k, reortho=...) of matfree_amd")


def funm_vectors(matvec, V, num_matvecs, params, matfun_code, reortho="full"):
    """``|v| Q^T f(T) e_1`` for the rows of ``V`` (P, n) via ``slq_funm_vectors``."""
    op = operators.resolve(matvec)
    P, n = tuple(V.shape)
    k = int(num_matvecs)
    if k < 0 or k > n:
        raise ValueError(decomp._error_num_matvecs(k, maxval=n, minval=0))
    device = _runtime.target_device(V, *params)
    V = _runtime.as_f32(V, device).reshape(P, n)
    desc, keep = op.descriptor(n, params, device)
    out = torch.empty((P, n), dtype=torch.float32, device=device)
    ctx = _runtime.context(device)
    ctx.call("slq_funm_vectors", ctypes.byref(desc), _runtime.ptr(V), P, k,
             1 if reortho == "full" else 0, int(matfun_code), _runtime.ptr(out))
    del keep
    return out


class _FunmLanczosSym:
    """``funm_lanczos_sym(dense_funm, tridiag_sym)``'s ``estimate`` (funm.py:114-147)."""

    def __init__(self, dense_funm, tridiag_sym):
        if not isinstance(dense_funm, _DenseFunmSymEigh) or dense_funm.code is None:
            raise NotImplementedError(
                "matfree_amd's funm_lanczos_sym applies dense_funm_sym_eigh with log, exp, sin, "
                "cos, sqrt or reciprocal"
            )
        self.reortho = _reortho_of(tridiag_sym)
        self.code = dense_funm.code
        self.num_matvecs = tridiag_sym.num_matvecs

    def batched(self, matvec, vecs, *parameters):
        """``f(A) v`` for every row of ``vecs`` (P, n) in one call."""
        return funm_vectors(matvec, vecs, self.num_matvecs, parameters, self.code, self.reortho)

    def __call__(self, matvec, vec, *parameters):
        flat, unravel = _tree.ravel(vec, decomp._host_or_gpu(vec, *parameters))
        out = funm_vectors(matvec, flat[None, :], self.num_matvecs, parameters, self.code,
                           self.reortho)
        return unravel(out[0])


def funm_lanczos_sym(dense_funm, tridiag_sym, /):
    """Matrix-function-vector product via Lanczos' tridiagonalisation (funm.py:114-147)."""
    return _FunmLanczosSym(dense_funm, tridiag_sym)


def monte_carlo_funm_sym(dense_funm, tridiag_sym, /):
    """Integrand for matrix-function trace estimation (funm.py:205-243)."""
    return _QuadFormSym(dense_funm, tridiag_sym)


def monte_carlo_funm_sym_logdet(tridiag_sym, /):
    """Integrand for the log-determinant (funm.py:186-202)."""
    return monte_carlo_funm_sym(dense_funm_sym_eigh(np.log), tridiag_sym)


# The name used by newer matfree API generations.
integrand_funm_sym_logdet = monte_carlo_funm_sym_logdet
