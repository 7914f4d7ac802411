"""NumPy restatement of matfree's SLQ log-determinant path -- TEST INFRASTRUCTURE.

Follows, in the same operation order (per probe; vectorised over probes where
the reference uses ``vmap``):

* ``decomp._hessenberg_forward`` / ``_hessenberg_forward_step``
  (decomp.py:426-477) -- Arnoldi with classical Gram-Schmidt applied twice;
* ``decomp._tridiag_reortho_full`` + ``_todense_tridiag_sym``
  (decomp.py:125-152) -- ``T = (H + H^T) / 2`` and its two diagonals;
* ``decomp._tridiag_forward`` (decomp.py:198-266) -- plain three-term Lanczos
  (``reortho="none"``);
* ``funm.monte_carlo_funm_sym`` / ``monte_carlo_funm_sym_logdet`` /
  ``dense_funm_sym_eigh`` (funm.py:186-243, 322-335);
* ``stochtrace.estimator_monte_carlo`` (stochtrace.py:47-50) and
  ``estimator_monte_carlo_mean_and_sem`` (stochtrace.py:55-89);
* ``funm.funm_lanczos_sym`` with ``dense_funm_sym_eigh`` (funm.py:114-147);
* ``decomp._hessenberg_adjoint`` / ``_hessenberg_adjoint_step``
  (decomp.py:480-600) and the eigh VJP that JAX applies at funm.py:331;
* ``decomp._tridiag_adjoint`` / ``_tridiag_adjoint_step`` (decomp.py:268-348),
  the reverse pass of ``reortho="none"``.

``dtype`` selects float32 (the reference's default precision) or float64.
All arrays are NumPy; nothing here touches a GPU.
"""

import numpy as np

from . import prng


# ---------------------------------------------------------------------------
# Operators (the reference's ``matvec(v, *params)`` plugin, batched over rows)
# ---------------------------------------------------------------------------


class DenseOperator:
    """``matvec(v, A) = A @ v`` as in tests/test_decomp/test_tridiag_sym.py:16-18.

    Acts on a batch ``V`` of shape ``(P, n)`` row by row.
    """

    def __init__(self, A, dtype=np.float32):
        self.A = np.asarray(A, dtype=dtype)
        self.n = self.A.shape[0]
        self.dtype = dtype

    def apply(self, V):
        return (V @ self.A.T).astype(self.dtype, copy=False)

    def apply_transpose(self, V):
        return (V @ self.A).astype(self.dtype, copy=False)

    def param_vjp(self, lam, q):
        """Cotangent of ``A`` for ``<lam, A q>``: the outer product."""
        return np.outer(lam, q)

    def dense(self):
        return np.asarray(self.A, dtype=np.float64)


class RBFOperator:
    """``K(x; l) v + s2 v`` with ``K_ij = exp(-|x_i - x_j|^2 / (2 l^2))``.

    The matrix-free squared-exponential covariance of tutorials/8_gaussian_logpdf.py:21-31
    (there ``exp(-sq_dists / (2 ell**2)) + noise * eye``).  The matrix is
    assembled once here (the oracle runs at small n only).
    """

    def __init__(self, X, lengthscale, noise, dtype=np.float32):
        self.X = np.asarray(X, dtype=np.float64)
        self.n = self.X.shape[0]
        self.lengthscale = float(lengthscale)
        self.noise = float(noise)
        self.dtype = dtype
        d2 = np.sum((self.X[:, None, :] - self.X[None, :, :]) ** 2, axis=-1)
        self._d2 = d2
        self._K64 = np.exp(-d2 / (2.0 * self.lengthscale**2))
        self._A = (self._K64 + self.noise * np.eye(self.n)).astype(dtype)

    def apply(self, V):
        return (V @ self._A).astype(self.dtype, copy=False)

    apply_transpose = apply

    def param_vjp(self, lam, q):
        """Cotangents of ``(lengthscale, noise)`` for ``<lam, A(l, s2) q>``."""
        dK_dl = self._K64 * self._d2 / self.lengthscale**3
        return np.array([lam @ dK_dl @ q, lam @ q])

    def dense(self):
        return self._K64 + self.noise * np.eye(self.n)


# ---------------------------------------------------------------------------
# Forward: Arnoldi with full reorthogonalisation (decomp.py:426-477)
# ---------------------------------------------------------------------------


def _error_num_matvecs(num, maxval, minval):
    # Message format of decomp.py:753-756 (tests match on "exceeds").
    return (
        f"Parameter 'num_matvecs'={num} exceeds the acceptable range. "
        f"Expected: {minval} <= num_matvecs <= {maxval}."
    )


def _hessenberg_step(Q, H, v, length, apply, idx):
    """One step of decomp.py:454-477, batched over the leading probe axis.

    Q: (P, n, k), H: (P, k, k), v: (P, n), length: (P,)
    """
    k = Q.shape[2]
    v = v / length[:, None]  # :456
    Q[:, :, idx] = v  # :457
    v = apply(v)  # :460
    h = np.matmul(np.swapaxes(Q, 1, 2), v[:, :, None])[:, :, 0]  # :463 h = Q^T v
    v = v - np.matmul(Q, h[:, :, None])[:, :, 0]  # :464
    h2 = np.matmul(np.swapaxes(Q, 1, 2), v[:, :, None])  # :468 (coefficients discarded)
    v = v - np.matmul(Q, h2)[:, :, 0]
    length = np.sqrt(np.sum(v * v, axis=1))  # :471 sqrt(inner(v, v))
    if idx + 1 < k:  # :474 (JAX drops the out-of-range update at idx + 1 = k)
        h[:, idx + 1] = length
    H[:, :, idx] = h  # :475
    return Q, H, v, length


def hessenberg_forward(operator, num_matvecs, V, dtype=np.float32):
    """``_hessenberg_forward`` (decomp.py:426-451) for a batch of start vectors.

    Returns ``(Q, H, r, c)`` with Q of shape (P, n, k) (columns = Krylov
    vectors), H (P, k, k), residual r (P, n) and ``c = 1 / |v|`` (P,).
    """
    V = np.asarray(V, dtype=dtype)
    if V.ndim == 1:
        V = V[None, :]
    P, n = V.shape
    k = int(num_matvecs)
    if k < 0 or k > n:  # :427-429
        raise ValueError(_error_num_matvecs(k, maxval=n, minval=0))
    Q = np.zeros((P, n, k), dtype=dtype)
    H = np.zeros((P, k, k), dtype=dtype)
    initlength = np.sqrt(np.sum(V * V, axis=1))  # :435
    v, length = V, initlength
    for i in range(k):  # :448 fori_loop
        Q, H, v, length = _hessenberg_step(Q, H, v, length, operator.apply, i)
    return Q, H, v, (1 / initlength).astype(dtype)


def tridiag_sym(operator, num_matvecs, V, dtype=np.float32, materialize=True):
    """``decomp.tridiag_sym(k, reortho="full")`` (decomp.py:125-152), batched.

    Returns ``(Q_tall, J_small, residual, init_length_inv)`` where Q_tall is
    (P, k, n) (rows = Krylov vectors, decomp.py:388) and, as in the reference
    (decomp.py:131, 142), the fourth entry is ``1 / c = |v|``.
    """
    Q, H, r, c = hessenberg_forward(operator, num_matvecs, V, dtype=dtype)
    T = 0.5 * (H + np.swapaxes(H, 1, 2))  # :133
    diags = np.diagonal(T, 0, axis1=1, axis2=2).copy()  # :134
    offdiags = np.diagonal(T, 1, axis1=1, axis2=2).copy()  # :135
    if materialize:  # :138-139, :148-152
        J = np.zeros_like(T)
        k = T.shape[1]
        idx = np.arange(k)
        J[:, idx, idx] = diags
        J[:, idx[:-1], idx[:-1] + 1] = offdiags
        J[:, idx[:-1] + 1, idx[:-1]] = offdiags
    else:
        J = (diags, offdiags)
    return np.swapaxes(Q, 1, 2), J, r, (1.0 / c).astype(dtype)


# ---------------------------------------------------------------------------
# SLQ integrand and estimator (funm.py:186-243, 322-335; stochtrace.py:47-50)
# ---------------------------------------------------------------------------


def lanczos_forward(operator, num_matvecs, V, dtype=np.float32):
    """``_tridiag_forward`` (decomp.py:198-266): plain three-term Lanczos, batched.

    Returns ``(Q (P, k, n), diags (P, k), offdiags (P, k-1), (q_k (P, n),
    b_k (P,)), c)`` with ``c = 1 / |v|`` (decomp.py:231).  Per step
    ``a = v . A v``, ``r = A v - a v - b v_prev``, ``b = |r|``, ``x = r / b``
    (decomp.py:243-247, 254-262).
    """
    V = np.asarray(V, dtype=dtype)
    P, n = V.shape
    k = num_matvecs
    if k < 0 or k > n:
        raise ValueError(_error_num_matvecs(k, maxval=n, minval=0))
    length = np.sqrt(np.sum(V * V, axis=1))
    vectors = np.zeros((k + 1, P, n), dtype=dtype)  # :201
    offdiags = np.zeros((P, k), dtype=dtype)
    diags = np.zeros((P, k), dtype=dtype)
    v0 = V / length[:, None]  # :208
    vectors[0] = v0

    def step(vec, b, vec_prev):
        w = operator.apply(vec).astype(dtype)
        a = np.sum(vec * w, axis=1).astype(dtype)
        r = w - a[:, None] * vec
        if vec_prev is not None:
            r = r - b[:, None] * vec_prev
        bn = np.sqrt(np.sum(r * r, axis=1)).astype(dtype)
        return r / bn[:, None], bn, a

    x, b, a = step(v0, None, None)  # _tridiag_fwd_init (:237-247)
    if k == 0:
        return vectors[:0].swapaxes(0, 1), diags, offdiags[:, :0], (x, b), (1 / length).astype(dtype)
    vectors[1], offdiags[:, 0], diags[:, 0] = x, b, a
    prev = v0
    for i in range(1, k):  # :225-228, _tridiag_fwd_step
        cur = vectors[i]
        x, b, a = step(cur, offdiags[:, i - 1], prev)
        vectors[i + 1], offdiags[:, i], diags[:, i] = x, b, a
        prev = cur
    return (vectors[:-1].swapaxes(0, 1), diags, offdiags[:, :-1], (vectors[-1], offdiags[:, -1]),
            (1 / length).astype(dtype))


def quadform_logdet_lanczos(operator, num_matvecs, Z, dtype=np.float32):
    """``quadform`` with ``tridiag_sym(k, reortho="none")`` (funm.py:226-241)."""
    Z = np.asarray(Z, dtype=dtype)
    length = np.sqrt(np.sum(Z * Z, axis=1))
    V0 = Z / length[:, None]
    _, d, e, _, _ = lanczos_forward(operator, num_matvecs, V0, dtype=dtype)
    k = d.shape[1]
    J = np.zeros((Z.shape[0], k, k), dtype=dtype)
    idx = np.arange(k)
    J[:, idx, idx] = d
    J[:, idx[:-1], idx[:-1] + 1] = e
    J[:, idx[:-1] + 1, idx[:-1]] = e
    fA = dense_funm_sym_eigh_log(J)
    return length**2 * fA[:, 0, 0], d, e


def dense_funm_sym_eigh_log(T):
    """``dense_funm_sym_eigh(np.log)`` (funm.py:322-335) on a batch of (k, k)."""
    w, U = np.linalg.eigh(T)  # :331 (LAPACK syevd; float32 stays float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        fw = np.log(w)  # :332 (no guard: NaN for negative Ritz values)
    return np.matmul(U * fw[:, None, :], np.swapaxes(U, 1, 2))  # :333


def quadform_logdet(operator, num_matvecs, Z, dtype=np.float32):
    """``monte_carlo_funm_sym(...).quadform`` (funm.py:226-241), batched.

    Returns the per-probe values ``|z|^2 [log T]_00`` (P,) and the tridiagonal
    coefficients ``(alpha (P, k), beta (P, k-1))`` for parity checks.
    """
    Z = np.asarray(Z, dtype=dtype)
    length = np.sqrt(np.sum(Z * Z, axis=1))  # :228 vector_norm
    V0 = Z / length[:, None]  # :229
    _, J, _, _ = tridiag_sym(operator, num_matvecs, V0, dtype=dtype)  # :237
    fA = dense_funm_sym_eigh_log(J)  # :239
    values = length**2 * fA[:, 0, 0]  # :240-241
    k = J.shape[1]
    alpha = np.diagonal(J, 0, axis1=1, axis2=2).copy()
    beta = np.diagonal(J, 1, axis1=1, axis2=2).copy() if k > 1 else np.zeros((len(Z), 0), dtype)
    return values.astype(dtype), alpha, beta


def estimator_mean(values, dtype=np.float32):
    """``tree_map(lambda s: mean(s, axis=0), Qs)`` (stochtrace.py:50)."""
    return np.mean(np.asarray(values, dtype=dtype), axis=0, dtype=dtype)


def estimator_mean_and_sem(values, dtype=np.float32):
    """``estimator_monte_carlo_mean_and_sem`` reductions (stochtrace.py:84-88):
    ``mean(s)`` and ``std(s) / sqrt(num_samples)`` (population std, ddof 0)."""
    s = np.asarray(values, dtype=dtype)
    return np.mean(s, axis=0, dtype=dtype), (np.std(s, axis=0, dtype=dtype)
                                             / np.sqrt(s.shape[0])).astype(dtype)


def dense_funm_sym_eigh(matfun, T):
    """``dense_funm_sym_eigh(matfun)`` (funm.py:322-335) on a batch of (k, k)."""
    w, U = np.linalg.eigh(T)  # :331
    with np.errstate(invalid="ignore", divide="ignore"):
        fw = matfun(w)  # :332
    return np.matmul(U * fw[:, None, :], np.swapaxes(U, 1, 2))  # :333


def funm_lanczos_sym(operator, num_matvecs, V, matfun, dtype=np.float32, reortho="full"):
    """``funm_lanczos_sym(dense_funm_sym_eigh(matfun), tridiag_sym(k, reortho))``
    (funm.py:114-147) for the rows of ``V`` (P, n): ``|v| Q^T f(T) e_1`` per row."""
    V = np.asarray(V, dtype=dtype)
    length = np.sqrt(np.sum(V * V, axis=1))  # :138
    V0 = V / length[:, None]  # :139
    if reortho == "full":
        Q, J, _, _ = tridiag_sym(operator, num_matvecs, V0, dtype=dtype)  # :144, Q (P, k, n)
    else:
        Q, d, e, _, _ = lanczos_forward(operator, num_matvecs, V0, dtype=dtype)
        k = d.shape[1]
        J = np.zeros((V.shape[0], k, k), dtype=dtype)
        idx = np.arange(k)
        J[:, idx, idx] = d
        J[:, idx[:-1], idx[:-1] + 1] = e
        J[:, idx[:-1] + 1, idx[:-1]] = e
    fJ = dense_funm_sym_eigh(matfun, J)  # :145
    y = fJ[:, :, 0]  # funm @ e1 (:146-147)
    return length[:, None] * np.einsum("pkn,pk->pn", Q, y)  # :147 Q.T @ (funm @ e1)


def slq_logdet(operator, num_matvecs, key, num_probes, dtype=np.float32):
    """Full estimator: Rademacher probes -> tridiag_sym -> logdet integrand -> mean.

    The composition of tutorials/1_log_determinants.py:29-35 /
    stochtrace.py:47-50 with ``sampler_signs`` (stochtrace.py:932-937).
    """
    Z = prng.rademacher(key, (num_probes, operator.n), dtype)
    values, alpha, beta = quadform_logdet(operator, num_matvecs, Z, dtype=dtype)
    return estimator_mean(values, dtype=dtype), values, alpha, beta


# ---------------------------------------------------------------------------
# FP64 tridiagonal eigen-solve + Gauss quadrature (the GPU kernel's checker)
# ---------------------------------------------------------------------------


def tridiag_eig_quadrature(alpha, beta, length):
    """``length^2 * sum_j U_0j^2 log(lambda_j)`` for the tridiagonal (alpha, beta).

    This is ``|z|^2 [log T]_00`` of funm.py:239-241 (``length = |z|``)
    evaluated in float64 via ``numpy.linalg.eigh`` of the dense tridiagonal.
    """
    alpha = np.asarray(alpha, np.float64)
    beta = np.asarray(beta, np.float64)
    out = np.empty(alpha.shape[0])
    for p in range(alpha.shape[0]):
        T = np.diag(alpha[p]) + np.diag(beta[p], 1) + np.diag(beta[p], -1)
        w, U = np.linalg.eigh(T)
        with np.errstate(invalid="ignore", divide="ignore"):
            out[p] = float(length[p]) ** 2 * np.sum(U[0] ** 2 * np.log(w))
    return out


def logdet_cholesky(A):
    """FP64 dense Cholesky log-determinant (the accuracy reference)."""
    L = np.linalg.cholesky(np.asarray(A, np.float64))
    return 2.0 * np.sum(np.log(np.diagonal(L)))


# ---------------------------------------------------------------------------
# Reverse mode: eigh VJP + T assembly + Arnoldi adjoint (decomp.py:480-600)
# ---------------------------------------------------------------------------


def _lower(m):
    """``lower`` of decomp.py:490-492."""
    m_tril = np.tril(m)
    return m_tril - 0.5 * np.diag(np.diag(m_tril))


def hessenberg_adjoint(operator, Q, H, r, c, dQ, dH, dr, dc, reortho="full"):
    """``_hessenberg_adjoint`` (decomp.py:480-600), one probe.

    Q: (n, k) columns = Krylov vectors; H: (k, k); r: (n,); c: scalar.
    ``reortho="none"`` skips the reorthogonalisation of the adjoint state
    (decomp.py:575).  Returns ``(dv, dp)`` where dp is the accumulated
    parameter cotangent of ``operator.param_vjp``.
    """
    n, k = Q.shape
    if k == 0:  # :483-486
        raise ValueError("Custom Hessenberg-adjoints are not implemented for num_matvecs = 0.")
    eye = np.eye(k)
    e_1, e_K = eye[0], eye[-1]  # :494
    lower_mask = _lower(np.ones((k, k)))  # :495

    gamma = dH @ e_K - Q.T @ dr  # :498
    lambda_k = dr + Q @ gamma  # :499
    Lambda = np.zeros_like(Q)  # :500
    Gamma = np.zeros((k, k))  # :501
    dp = None  # :502

    Pi_xi = dQ.T + np.outer(gamma, r)  # :505
    Pi_gamma = -dc * c * np.outer(e_1, e_1) + H @ dH.T - (dQ.T @ Q)  # :506
    reortho_mask = np.tril(np.ones((k, k)), 1)  # :509

    beta_minuses = np.concatenate([np.ones((1,)), np.diag(H, -1)])  # :513
    alphas = np.diag(H)  # :514
    beta_pluses = H - np.diag(np.diag(H)) - np.diag(np.diag(H, -1), -1)  # :515

    for idx in reversed(range(k)):  # :538 reverse scan
        # :576-584 reorthogonalisation of the adjoint state
        if reortho == "full":
            Q_masked = reortho_mask[idx][None, :] * Q
            rhs_masked = reortho_mask[idx] * dH.T[idx]
            lambda_k = lambda_k - Q_masked @ (Q_masked.T @ lambda_k) + Q_masked @ rhs_masked
        # :587-589 transposed matvec and parameter cotangent
        q = Q[:, idx]
        vecmat_lambda = operator.apply_transpose(lambda_k[None, :].astype(operator.dtype))[0]
        vecmat_lambda = np.asarray(vecmat_lambda, np.float64)
        dp_inc = operator.param_vjp(lambda_k, q)
        dp = dp_inc if dp is None else dp + dp_inc
        # :592-593
        tmp = lower_mask[idx] * (Pi_gamma[idx] - vecmat_lambda @ Q)
        Gamma[idx, :] = tmp
        # :596-599
        Lambda[:, idx] = lambda_k
        xi = Pi_xi[idx] + (Gamma + Gamma.T)[idx, :] @ Q.T
        lambda_k = xi - (alphas[idx] * lambda_k - vecmat_lambda) - beta_pluses[idx] @ Lambda.T
        lambda_k = lambda_k / beta_minuses[idx]
    dv = lambda_k * c  # :542
    return dv, dp


def logdet_quadform_cotangent_T(J):
    """Cotangent of ``[log J]_00`` w.r.t. the symmetric input of ``eigh``.

    JAX's eigh VJP (funm.py:331) composed with ``fA[0, 0]`` (funm.py:333, 241)
    and the input symmetrisation of ``jnp.linalg.eigh`` reduces to the
    Daleckii-Krein form ``U (L o u u^T) U^T`` with ``u = U[0, :]`` and ``L`` the
    divided differences of ``log`` at the eigenvalues.
    """
    w, U = np.linalg.eigh(J)
    u = U[0]
    k = len(w)
    L = np.empty((k, k))
    for i in range(k):
        for j in range(k):
            if i == j or w[i] == w[j]:
                L[i, j] = 1.0 / w[i]
            else:
                L[i, j] = (np.log(w[j]) - np.log(w[i])) / (w[j] - w[i])
    return U @ (L * np.outer(u, u)) @ U.T


def slq_logdet_vjp(operator, num_matvecs, Z, dtype=np.float64):
    """Gradient of ``mean_p |z_p|^2 [log T_p]_00`` w.r.t. the operator parameters.

    Chain: eigh VJP (funm.py:331) -> dense tridiagonal assembly
    (decomp.py:148-152) -> ``T = (H + H^T)/2`` diagonals (decomp.py:133-135)
    -> Arnoldi adjoint with ``dQ = dr = dc = 0`` (quadform discards them,
    funm.py:237) -> mean over probes (stochtrace.py:50).
    Returns ``(estimate, dparams)``.
    """
    Z = np.asarray(Z, dtype=dtype)
    P = Z.shape[0]
    length = np.sqrt(np.sum(Z * Z, axis=1))
    V0 = Z / length[:, None]
    Q, H, r, c = hessenberg_forward(operator, num_matvecs, V0, dtype=dtype)
    k = num_matvecs
    values = np.empty(P)
    grad = None
    for p in range(P):
        Hp = np.asarray(H[p], np.float64)
        T = 0.5 * (Hp + Hp.T)
        J = np.diag(np.diag(T)) + np.diag(np.diag(T, 1), 1) + np.diag(np.diag(T, 1), -1)
        w, U = np.linalg.eigh(J)
        values[p] = length[p] ** 2 * np.sum(U[0] ** 2 * np.log(w))
        Jbar = logdet_quadform_cotangent_T(J) * (length[p] ** 2 / P)
        # J = diag(d) + diag(e, 1) + diag(e, -1):  d_bar = diag(Jbar), e_bar = 2 diag(Jbar, 1)
        d_bar = np.diag(Jbar)
        e_bar = np.diag(Jbar, 1) + np.diag(Jbar, -1)
        # d = diag(H), e = (H[i, i+1] + H[i+1, i]) / 2
        dH = np.diag(d_bar) + 0.5 * np.diag(e_bar, 1) + 0.5 * np.diag(e_bar, -1)
        n = Q.shape[1]
        _, dp = hessenberg_adjoint(
            operator,
            np.asarray(Q[p], np.float64),
            Hp,
            np.asarray(r[p], np.float64),
            float(c[p]),
            dQ=np.zeros((n, k)),
            dH=dH,
            dr=np.zeros(n),
            dc=0.0,
        )
        grad = dp if grad is None else grad + dp
    return np.mean(values), grad


# ---------------------------------------------------------------------------
# Reverse: plain Lanczos adjoint (decomp.py:268-348, reortho="none")
# ---------------------------------------------------------------------------


def _lanczos_adjoint_step(operator, xi, lam_plus, x, xplus, a, b, da, db, dx):
    """One reverse step of ``_tridiag_adjoint_step`` (decomp.py:321-348)."""
    xi = xi / b
    mu = db - lam_plus @ x + xplus @ xi
    nu = da + x @ xi
    lam = -xi + mu * xplus + nu * x
    Alam = operator.apply(lam[None])[0]
    g = operator.param_vjp(x, lam)  # vjp of matvec(lambda) at cotangent x (:337-338)
    xi = -dx - Alam + a * lam + b * lam_plus - b * nu * xplus
    return xi, lam, g


def lanczos_adjoint(operator, alphas, betas, xs, dalphas, dbetas, dxs, initvec_norm):
    """``_tridiag_adjoint`` + ``_tridiag_adjoint_step`` (decomp.py:268-348) for one probe.

    ``xs`` (k + 1, n) are the Lanczos vectors including the remainder's
    direction, ``betas`` (k,) the off-diagonals including the remainder's
    norm, ``dalphas``/``dbetas`` (k,) and ``dxs`` (k + 1, n) their cotangents
    (decomp.py:190-197 stacks them the same way).  The reverse scan runs
    i = k-1 .. 0 from ``xi = -dxs[-1]``, ``lambda_plus = 0`` (:290-297):

        xi /= b_i;  mu = db_i - lambda_plus . x_i + x_{i+1} . xi
        nu = da_i + x_i . xi;  lambda = -xi + mu x_{i+1} + nu x_i
        grad += d/dp <x_i, A(p) lambda>      (vjp of matvec(lambda) at x_i)
        xi = -dx_i - A lambda + a_i lambda + b_i lambda_plus - b_i nu x_{i+1}

    Returns ``(grad_initvec, grad_params, lambdas)`` with
    ``grad_initvec = ((xi . x_0) x_0 - xi) / |v|`` for the final ``xi``
    (:298-305: the scan's carry is ``(xs, xi, lambda)``, so its second entry,
    named ``lambda_1`` there, is ``xi``).
    """
    k = len(alphas)
    xi = -np.asarray(dxs[-1], np.float64)
    lam_plus = np.zeros_like(xi)
    grad = None
    lambdas = np.empty((k, xi.shape[0]))
    for i in range(k - 1, -1, -1):
        xi, lam, g = _lanczos_adjoint_step(operator, xi, lam_plus, xs[i], xs[i + 1], alphas[i],
                                           betas[i], dalphas[i], dbetas[i], dxs[i])
        grad = g if grad is None else grad + g
        lam_plus = lam
        lambdas[i] = lam
    grad_initvec = ((xi @ xs[0]) * xs[0] - xi) / initvec_norm
    return grad_initvec, grad, lambdas


def slq_logdet_lanczos_vjp(operator, num_matvecs, Z, dtype=np.float64):
    """Gradient of ``mean_p |z_p|^2 [log T_p]_00`` with ``tridiag_sym(reortho="none")``.

    Chain: eigh VJP (funm.py:331) -> dense tridiagonal assembly
    (decomp.py:148-152: ``d_bar = diag(Jbar)``, ``e_bar = 2 diag(Jbar, 1)``)
    -> Lanczos adjoint with ``dxs = 0`` and a zero cotangent for the
    remainder's norm (quadform discards Q and the residual, funm.py:237)
    -> mean over probes (stochtrace.py:50).  Returns ``(estimate, dparams)``.
    """
    Z = np.asarray(Z, dtype=dtype)
    P, n = Z.shape
    k = num_matvecs
    if k < 1:
        raise ValueError("the log-determinant integrand needs num_matvecs >= 1")
    length = np.sqrt(np.sum(Z * Z, axis=1))
    V0 = Z / length[:, None]
    Q, d, e, (qk, bk), _ = lanczos_forward(operator, k, V0, dtype=dtype)
    values = np.empty(P)
    grad = None
    for p in range(P):
        dp_, ep = np.asarray(d[p], np.float64), np.asarray(e[p], np.float64)
        J = np.diag(dp_) + np.diag(ep, 1) + np.diag(ep, -1)
        w, U = np.linalg.eigh(J)
        values[p] = length[p] ** 2 * np.sum(U[0] ** 2 * np.log(w))
        Jbar = logdet_quadform_cotangent_T(J) * (length[p] ** 2 / P)
        dalphas = np.diag(Jbar).copy()
        dbetas = np.zeros(k)
        dbetas[: k - 1] = np.diag(Jbar, 1) + np.diag(Jbar, -1)
        xs = np.concatenate([np.asarray(Q[p], np.float64), np.asarray(qk[p], np.float64)[None]])
        betas = np.concatenate([ep, [float(bk[p])]])
        _, g, _ = lanczos_adjoint(operator, dp_, betas, xs, dalphas, dbetas,
                                  np.zeros((k + 1, n)), float(length[p]))
        grad = g if grad is None else grad + g
    return np.mean(values), grad
