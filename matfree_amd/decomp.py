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


def arnoldi(matvec, V, k, params):
    """Run ``slq_hessenberg`` on the rows of ``V`` (P, n).

    Returns ``Q`` (k, P, n), ``H`` (P, k, k) row-major, ``R`` (P, n), ``c`` (P,).
    """
    op = operators.resolve(matvec)
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
        "slq_hessenberg",
        ctypes.byref(desc),
        _runtime.ptr(V),
        P,
        k,
        _runtime.ptr(Q) if k > 0 else None,
        _runtime.ptr(Hc) if k > 0 else None,
        _runtime.ptr(R),
        _runtime.ptr(c),
    )
    del keep
    H = Hc.transpose(1, 2)  # column-major per probe -> (row, col)
    return Q, H, R, c


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
    ``c = 1 / |v|`` (P,).
    """
    op = operators.resolve(matvec)
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

    def __init__(self, num_matvecs, custom_vjp, reortho_vjp):
        self.num_matvecs = int(num_matvecs)
        self.custom_vjp = bool(custom_vjp)
        self.reortho_vjp = reortho_vjp

    def batched(self, matvec, vecs, *params):
        Q, H, R, c = arnoldi(matvec, vecs, self.num_matvecs, params)
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
    (decomp.py:394-396, 467): with the default ``reortho_vjp="match"`` both
    ``reortho="full"`` and ``reortho="none"`` run the full-reorthogonalisation
    kernels.  ``reortho_vjp="none"`` raises ``NotImplementedError``.
    """
    reortho_expected = ["none", "full"]
    if reortho not in reortho_expected:  # decomp.py:375-378
        msg = f"Unexpected input for {reortho}: either of {reortho_expected} expected."
        raise TypeError(msg)
    if reortho_vjp == "none":
        raise NotImplementedError("hessenberg forward without reorthogonalisation is not provided")
    return _Hessenberg(num_matvecs, custom_vjp=custom_vjp, reortho_vjp=reortho_vjp)
