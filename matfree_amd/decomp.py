"""Matrix decompositions on the GPU: ``tridiag_sym`` and ``hessenberg``.

Drop-in for matfree's ``decomp.tridiag_sym(num_matvecs, *, materialize,
reortho, custom_vjp)`` (decomp.py:30-122) and ``decomp.hessenberg``
(decomp.py:351-423) with full reorthogonalisation.  The factories keep the
reference's signatures, validation and error messages; the returned
``decompose(matvec, vec, *params)`` runs libslq's Arnoldi kernels
(``slq_hessenberg``) and returns ``_DecompResult(Q_tall, J_small, residual,
init_length_inv)`` as torch tensors on the GPU.

Every decomposition also has a ``batched(matvec, vecs, *params)`` form that
factorises P start vectors (rows of ``vecs``) at once -- the role ``vmap``
plays in the reference (stochtrace.py:49).
"""

import ctypes
from typing import Any, NamedTuple

import torch

from . import _runtime, _tree, operators


class _DecompResult(NamedTuple):
    """Same fields as decomp.py:15-27."""

    Q_tall: Any
    J_small: Any
    residual: Any
    init_length_inv: Any


def _error_num_matvecs(num, maxval, minval):
    # decomp.py:753-756
    msg1 = f"Parameter 'num_matvecs'={num} exceeds the acceptable range. "
    msg2 = f"Expected: {minval} <= num_matvecs <= {maxval}."
    return msg1 + msg2


def _arnoldi_raw(op, V, k, params, reortho=True):
    P, n = tuple(V.shape)
    if k < 0 or k > n:  # decomp.py:427-429
        raise ValueError(_error_num_matvecs(k, maxval=n, minval=0))
    device = _runtime.target_device(V, *params)
    V = _runtime.as_f32(V, device)
    desc, keep = op.descriptor(n, params, device)
    Q = torch.empty((k, P, n), dtype=torch.float32, device=device)
    Hc = torch.empty((P, k, k), dtype=torch.float32, device=device)
    R = torch.empty((P, n), dtype=torch.float32, device=device)
    c = torch.empty((P,), dtype=torch.float32, device=device)
    ctx = _runtime.context(device)
    ctx.call(
        "slq_arnoldi",
        ctypes.byref(desc),
        _runtime.ptr(V),
        P,
        k,
        1 if reortho else 0,
        _runtime.ptr(Q) if k > 0 else None,
        _runtime.ptr(Hc) if k > 0 else None,
        _runtime.ptr(R),
        _runtime.ptr(c),
    )
    del keep
    H = Hc.transpose(1, 2)  # column-major per probe -> (row, col)
    return Q, H, R, c


def _param_grads(op, params, dparams, dmatrix, needs):
    """Cotangents of the operator parameters in the order of ``params``
    (None where a parameter is not a tensor or needs no gradient)."""
    out = []
    for i, p in enumerate(params):
        if not (isinstance(p, torch.Tensor) and needs[i]):
            out.append(None)
            continue
        if isinstance(op, operators.RBF):
            g = dparams[i]
        else:
            g = dmatrix
        out.append(g.to(dtype=p.dtype, device=p.device).reshape(p.shape))
    return out


def _grad_buffers(op, n, device):
    if isinstance(op, operators.RBF):
        return torch.zeros((2,), dtype=torch.float64, device=device), None
    return None, torch.zeros((n, n), dtype=torch.float32, device=device)


class _ArnoldiFn(torch.autograd.Function):
    """Arnoldi forward (``slq_arnoldi``) with the reference's custom VJP
    (decomp.py:398-422): backward runs ``slq_arnoldi_vjp`` on the incoming
    cotangents of (Q, H, r, c)."""

    @staticmethod
    def forward(ctx, matvec, k, reortho_fwd, reortho_adj, V, *params):
        op = operators.resolve(matvec)
        Q, H, R, c = _arnoldi_raw(op, V.detach(), k, params, reortho_fwd)
        ctx.op, ctx.k, ctx.reortho_fwd, ctx.reortho_adj = op, k, reortho_fwd, reortho_adj
        ctx.params = params
        ctx.save_for_backward(V)
        return Q, H.contiguous(), R, c

    @staticmethod
    def backward(ctx, dQ, dH, dR, dc):
        (V,) = ctx.saved_tensors
        op, k, params = ctx.op, ctx.k, ctx.params
        P, n = tuple(V.shape)
        device = _runtime.target_device(V, *params)
        V32 = _runtime.as_f32(V.detach(), device)
        desc, keep = op.descriptor(n, params, device)
        dparams, dmatrix = _grad_buffers(op, n, device)
        dv = torch.empty((P, n), dtype=torch.float32, device=device)
        f32 = (lambda t: _runtime.as_f32(t, device) if t is not None else None)
        dQc = f32(dQ)
        dHc = f32(dH.transpose(1, 2)) if dH is not None else None  # to column-major
        dRc, dcc = f32(dR), f32(dc)
        cx = _runtime.context(device)
        cx.call("slq_arnoldi_vjp", ctypes.byref(desc), _runtime.ptr(V32), P, k,
                1 if ctx.reortho_fwd else 0, 1 if ctx.reortho_adj else 0,
                _runtime.ptr(dQc), _runtime.ptr(dHc), _runtime.ptr(dRc), _runtime.ptr(dcc),
                _runtime.ptr(dv), _runtime.ptr(dparams), _runtime.ptr(dmatrix))
        del keep
        grads = _param_grads(op, params, dparams, dmatrix, ctx.needs_input_grad[5:])
        return (None, None, None, None, dv.to(V.dtype) if ctx.needs_input_grad[4] else None,
                *grads)


def _differentiable(V, params):
    return torch.is_grad_enabled() and (
        (isinstance(V, torch.Tensor) and V.requires_grad)
        or any(isinstance(p, torch.Tensor) and p.requires_grad for p in params))


def arnoldi(matvec, V, k, params, reortho=True, reortho_adjoint=True):
    """Run ``slq_arnoldi`` on the rows of ``V`` (P, n).

    Returns ``Q`` (k, P, n), ``H`` (P, k, k) row-major, ``R`` (P, n), ``c`` (P,).
    Differentiable in ``V`` and tensor parameters through the Arnoldi
    adjoint (``slq_arnoldi_vjp``) when any of them requires grad.
    """
    op = operators.resolve(matvec)
    if _differentiable(V, params):
        if not isinstance(V, torch.Tensor):
            V = torch.as_tensor(V)
        return _ArnoldiFn.apply(op, int(k), bool(reortho), bool(reortho_adjoint), V, *params)
    return _arnoldi_raw(op, V, int(k), params, reortho)


def _host_or_gpu(*candidates):
    """Where to flatten inputs: their GPU if they have one, else the host.

    Flattening on the host lets shape errors (e.g. ``num_matvecs`` out of
    range) surface before any device is required.
    """
    for c in candidates:
        if isinstance(c, torch.Tensor) and c.is_cuda:
            return c.device
    return torch.device("cpu")


def _symmetric_part(H):
    # T = 0.5 * (H + H^T), its diagonal and first off-diagonal (decomp.py:133-135)
    T = 0.5 * (H + H.transpose(-1, -2))
    diags = torch.diagonal(T, 0, -2, -1)
    offdiags = torch.diagonal(T, 1, -2, -1)
    return diags, offdiags


def _todense_tridiag_sym(diag, off_diag):
    # decomp.py:148-152
    return torch.diag_embed(diag) + torch.diag_embed(off_diag, -1) + torch.diag_embed(off_diag, 1)


class _TridiagSymFull:
    """``decompose`` of ``tridiag_sym(k, reortho="full")`` (decomp.py:125-145)."""

    def __init__(self, num_matvecs, materialize, custom_vjp):
        self.num_matvecs = int(num_matvecs)
        self.materialize = bool(materialize)
        self.custom_vjp = bool(custom_vjp)

    def batched(self, matvec, vecs, *params):
        Q, H, R, c = arnoldi(matvec, vecs, self.num_matvecs, params)
        diags, offdiags = _symmetric_part(H)
        matrix = _todense_tridiag_sym(diags, offdiags) if self.materialize else (diags, offdiags)
        Q_tall = Q.permute(1, 0, 2)  # (P, k, n): rows are Krylov vectors (decomp.py:388)
        # 4th field: 1 / (1 / |v|) = |v| (decomp.py:131, 142 with :387-391)
        return _DecompResult(Q_tall=Q_tall, J_small=matrix, residual=R, init_length_inv=1.0 / c)

    def __call__(self, matvec, vec, *params):
        flat, unravel = _tree.ravel(vec, _host_or_gpu(vec, *params))
        out = self.batched(matvec, flat[None, :], *params)
        matrix = out.J_small[0] if self.materialize else (out.J_small[0][0], out.J_small[1][0])
        return _DecompResult(
            Q_tall=unravel(out.Q_tall[0]),
            J_small=matrix,
            residual=unravel(out.residual[0]),
            init_length_inv=out.init_length_inv[0],
        )


def lanczos(matvec, V, k, params):
    """Run ``slq_lanczos`` (plain three-term Lanczos) on the rows of ``V`` (P, n).

    Returns ``Q`` (k, P, n), ``alpha`` (P, k), ``beta`` (P, max(k, 1)) -- the
    off-diagonals of T, then the remainder's norm --, ``R`` (P, n) and
    ``c = 1 / |v|`` (P,).  Differentiable in ``V`` and tensor parameters
    through the Lanczos adjoint (``slq_lanczos_vjp``) when any of them
    requires grad (the remainder enters as ``R = b x_k``, 1 / |v| directly
    in ``V``, as in decomp.py:164-178).
    """
    op = operators.resolve(matvec)
    if _differentiable(V, params) and int(k) > 0:
        if not isinstance(V, torch.Tensor):
            V = torch.as_tensor(V)
        Q, alpha, beta, xk = _LanczosFn.apply(op, int(k), V, *params)
        R = beta[:, -1:] * xk
        c = 1.0 / torch.linalg.vector_norm(V.to(torch.float32), dim=1)
        return Q, alpha, beta, R, c
    return _lanczos_raw(op, V, int(k), params)


def _lanczos_raw(op, V, k, params):
    P, n = tuple(V.shape)
    if k < 0 or k > n:  # decomp.py:160-162
        raise ValueError(_error_num_matvecs(k, maxval=n, minval=0))
    device = _runtime.target_device(V, *params)
    V = _runtime.as_f32(V, device)
    desc, keep = op.descriptor(n, params, device)
    Q = torch.empty((k, P, n), dtype=torch.float32, device=device)
    alpha = torch.empty((P, k), dtype=torch.float32, device=device)
    beta = torch.empty((P, max(k, 1)), dtype=torch.float32, device=device)
    R = torch.empty((P, n), dtype=torch.float32, device=device)
    c = torch.empty((P,), dtype=torch.float32, device=device)
    ctx = _runtime.context(device)
    ctx.call(
        "slq_lanczos",
        ctypes.byref(desc),
        _runtime.ptr(V),
        P,
        k,
        _runtime.ptr(Q) if k > 0 else None,
        _runtime.ptr(alpha) if k > 0 else None,
        _runtime.ptr(beta),
        _runtime.ptr(R),
        _runtime.ptr(c),
    )
    del keep
    return Q, alpha, beta, R, c


class _LanczosFn(torch.autograd.Function):
    """Plain Lanczos forward with the reference's custom VJP
    (decomp.py:181-197): outputs (Q, alpha, beta, x_k) with beta holding the
    remainder's norm last; backward runs ``slq_lanczos_vjp``."""

    @staticmethod
    def forward(ctx, matvec, k, V, *params):
        op = operators.resolve(matvec)
        Q, alpha, beta, R, _ = _lanczos_raw(op, V.detach(), k, params)
        xk = R / beta[:, -1:]
        ctx.op, ctx.k, ctx.params = op, k, params
        ctx.save_for_backward(V)
        return Q, alpha, beta, xk

    @staticmethod
    def backward(ctx, dQ, dalpha, dbeta, dxk):
        (V,) = ctx.saved_tensors
        op, k, params = ctx.op, ctx.k, ctx.params
        P, n = tuple(V.shape)
        device = _runtime.target_device(V, *params)
        V32 = _runtime.as_f32(V.detach(), device)
        desc, keep = op.descriptor(n, params, device)
        dparams, dmatrix = _grad_buffers(op, n, device)
        dv = torch.empty((P, n), dtype=torch.float32, device=device)
        f32 = (lambda t: _runtime.as_f32(t, device) if t is not None else None)
        dQc, dac, dbc, dxc = f32(dQ), f32(dalpha), f32(dbeta), f32(dxk)
        cx = _runtime.context(device)
        cx.call("slq_lanczos_vjp", ctypes.byref(desc), _runtime.ptr(V32), P, k,
                _runtime.ptr(dQc), _runtime.ptr(dac), _runtime.ptr(dbc), _runtime.ptr(dxc),
                _runtime.ptr(dv), _runtime.ptr(dparams), _runtime.ptr(dmatrix))
        del keep
        grads = _param_grads(op, params, dparams, dmatrix, ctx.needs_input_grad[3:])
        return (None, None, dv.to(V.dtype) if ctx.needs_input_grad[2] else None, *grads)


class _TridiagSymNone:
    """``decompose`` of ``tridiag_sym(k, reortho="none")`` (decomp.py:155-193)."""

    def __init__(self, num_matvecs, materialize, custom_vjp):
        self.num_matvecs = int(num_matvecs)
        self.materialize = bool(materialize)
        self.custom_vjp = bool(custom_vjp)

    def batched(self, matvec, vecs, *params):
        k = self.num_matvecs
        Q, alpha, beta, R, c = lanczos(matvec, vecs, k, params)
        offdiags = beta[:, : max(k - 1, 0)]
        matrix = _todense_tridiag_sym(alpha, offdiags) if self.materialize else (alpha, offdiags)
        # residual = b q (decomp.py:167); init_length_inv = 1 / |vec| (:171-178)
        return _DecompResult(Q_tall=Q.permute(1, 0, 2), J_small=matrix, residual=R,
                             init_length_inv=c)

    def __call__(self, matvec, vec, *params):
        flat, unravel = _tree.ravel(vec, _host_or_gpu(vec, *params))
        out = self.batched(matvec, flat[None, :], *params)
        matrix = out.J_small[0] if self.materialize else (out.J_small[0][0], out.J_small[1][0])
        return _DecompResult(
            Q_tall=unravel(out.Q_tall[0]),
            J_small=matrix,
            residual=unravel(out.residual[0]),
            init_length_inv=out.init_length_inv[0],
        )


def tridiag_sym(num_matvecs: int, /, *, materialize: bool = True, reortho: str = "full",
                custom_vjp: bool = True):
    """Tridiagonalisation of a symmetric operator (decomp.py:30-122).

    ``reortho="full"`` runs Arnoldi with classical Gram-Schmidt applied twice
    and symmetrises ``T = (H + H^T)/2`` (decomp.py:125-145); ``reortho="none"``
    runs plain three-term Lanczos (decomp.py:155-266).  Both on the GPU.
    """
    if reortho == "full":
        return _TridiagSymFull(num_matvecs, materialize=materialize, custom_vjp=custom_vjp)
    if reortho == "none":
        return _TridiagSymNone(num_matvecs, materialize=materialize, custom_vjp=custom_vjp)
    msg = f"reortho={reortho} unsupported. Choose eiter {'full', 'none'}."
    raise ValueError(msg)


class _Hessenberg:
    """``estimate`` of ``hessenberg(k, reortho=...)`` (decomp.py:380-391)."""

    def __init__(self, num_matvecs, reortho, custom_vjp, reortho_vjp):
        self.num_matvecs = int(num_matvecs)
        self.reortho = reortho
        self.custom_vjp = bool(custom_vjp)
        self.reortho_vjp = reortho_vjp

    def batched(self, matvec, vecs, *params):
        # the forward reorthogonalises unless reortho_vjp == "none"
        # (decomp.py:394-396, 467); the adjoint follows `reortho`
        # (decomp.py:413-420, 575)
        Q, H, R, c = arnoldi(matvec, vecs, self.num_matvecs, params,
                             reortho=self.reortho_vjp != "none",
                             reortho_adjoint=self.reortho == "full")
        return _DecompResult(Q_tall=Q.permute(1, 0, 2), J_small=H, residual=R, init_length_inv=c)

    def __call__(self, matvec, v, *params):
        flat, unravel = _tree.ravel(v, _host_or_gpu(v, *params))
        out = self.batched(matvec, flat[None, :], *params)
        return _DecompResult(
            Q_tall=unravel(out.Q_tall[0]),
            J_small=out.J_small[0],
            residual=unravel(out.residual[0]),
            init_length_inv=out.init_length_inv[0],
        )


def hessenberg(num_matvecs, /, *, reortho: str, custom_vjp: bool = True,
               reortho_vjp: str = "match"):
    """Hessenberg factorisation via Arnoldi (decomp.py:351-423).

    The forward pass reorthogonalises unless ``reortho_vjp == "none"``
    (decomp.py:394-396, 467: with the default ``reortho_vjp="match"`` both
    ``reortho="full"`` and ``reortho="none"`` run classical Gram-Schmidt
    twice); the reverse pass (the Arnoldi adjoint, decomp.py:480-600)
    reorthogonalises when ``reortho == "full"`` (decomp.py:575).  Outputs are
    differentiable in the start vector and tensor parameters when either
    requires grad; ``custom_vjp=False`` uses the same adjoint (the GPU loop
    is not traced).
    """
    reortho_expected = ["none", "full"]
    if reortho not in reortho_expected:  # decomp.py:375-378
        msg = f"Unexpected input for {reortho}: either of {reortho_expected} expected."
        raise TypeError(msg)
    return _Hessenberg(num_matvecs, reortho=reortho, custom_vjp=custom_vjp,
                       reortho_vjp=reortho_vjp)
