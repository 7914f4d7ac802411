"""Pins the oracle's SLQ restatement (a2-a8, a10-a11) by mathematics.

JAX is not installed here, so matfree itself cannot produce vectors; the
restatement is fixed instead by the reference's own invariants and by closed
forms chosen so that a dropped term, a wrong sign or index, or a transposed
operand fails at least one check (see the planted-mistake tests at the end).
"""

import numpy as np
import pytest

from oracle import prng, slq


def sym_from_eigs(eigvals, seed):
    # test_util.hermitian_matrix_from_eigenvalues (matfree/test_util.py:6-16), NumPy stream
    n = len(eigvals)
    Q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((n, n)))
    return (Q * eigvals) @ Q.T


# -- forward (decomp.py:426-477, 125-152) -------------------------------------


def test_full_rank_reconstruction_is_exact():
    # tests/test_decomp/test_tridiag_sym.py:7-38
    n = 12
    A = sym_from_eigs(np.arange(1.0, 2.0, 1 / n), 1)
    v = np.flip(np.arange(1.0, 1.0 + n))
    Qt, T, _, ln = slq.tridiag_sym(slq.DenseOperator(A), n, v[None])
    Q = Qt[0]
    assert np.allclose(Q.T @ T[0] @ Q, A, atol=1e-5, rtol=1e-5)
    assert np.allclose(Q @ Q.T, np.eye(n), atol=1e-5)
    assert np.allclose(Q.T @ Q, np.eye(n), atol=1e-5)
    assert np.isclose(ln[0], np.linalg.norm(v), rtol=1e-6)  # 4th field is |v| (decomp.py:131,142)


@pytest.mark.parametrize("k", [1, 5, 11])
def test_mid_rank_lanczos_relation(k):
    # tests/test_decomp/test_tridiag_sym.py:43-65: A Q^T = Q^T T + q e_K^T
    n = 12
    A = sym_from_eigs(np.arange(1.0, 2.0, 1 / n), 1)
    v = np.flip(np.arange(1.0, 1.0 + n))
    Qt, T, r, _ = slq.tridiag_sym(slq.DenseOperator(A), k, v[None])
    e_K = np.eye(k)[-1]
    tol = 10 * np.sqrt(np.finfo(np.float32).eps)
    assert np.allclose(A @ Qt[0].T - Qt[0].T @ T[0] - np.outer(r[0], e_K), 0.0, atol=tol)


@pytest.mark.parametrize("k", [0, 5, 9])
def test_arnoldi_relation_nonsymmetric(k):
    # tests/test_decomp/test_hessenberg.py:7-38 (general square A)
    n = 10
    rng = np.random.default_rng(3)
    A = rng.standard_normal((n, n))
    v = rng.standard_normal(n)
    Q, H, r, c = slq.hessenberg_forward(slq.DenseOperator(A, np.float64), k, v[None], np.float64)
    Q, H, r = Q[0], H[0], r[0]
    assert Q.shape == (n, k) and H.shape == (k, k) and r.shape == (n,)
    if k:
        e_k = np.eye(k)[-1]
        assert np.allclose(A @ Q - Q @ H - np.outer(r, e_k), 0.0, atol=1e-10)
        assert np.allclose(Q.T @ Q, np.eye(k), atol=1e-12)
        assert np.allclose(Q[:, 0], c[0] * v, atol=1e-12)


def test_rayleigh_ritz_identity():
    # T = Q A Q^T in exact arithmetic: pins every alpha and beta (index and sign)
    n, k = 40, 15
    A = sym_from_eigs(np.linspace(0.5, 3.0, n), 4)
    z = prng.rademacher(prng.prng_key(2), (3, n), np.float64)
    Qt, T, _, _ = slq.tridiag_sym(slq.DenseOperator(A, np.float64), k, z, np.float64)
    for p in range(3):
        assert np.allclose(Qt[p] @ A @ Qt[p].T, T[p], atol=1e-11)


def test_step_uses_first_pass_coefficients_only():
    # decomp.py:463-475: H[:, j] holds h = Q^T w from the first pass; the
    # second pass's coefficients are discarded.  With a non-orthonormal basis
    # the two differ at O(1), so this fixes which pass feeds H.
    rng = np.random.default_rng(5)
    n, k = 8, 4
    A = rng.standard_normal((n, n))
    Q = np.zeros((1, n, k))
    Q[0, :, 0] = rng.standard_normal(n)
    Q[0, :, 1] = rng.standard_normal(n)
    H = np.zeros((1, k, k))
    v = rng.standard_normal((1, n))
    length = np.array([2.0])
    Q2, H2, w, ln = slq._hessenberg_step(Q.copy(), H.copy(), v, length, lambda x: x @ A.T, 2)
    q2 = v[0] / 2.0
    Qf = Q2[0]
    w0 = A @ q2
    h1 = Qf.T @ w0
    w1 = w0 - Qf @ h1
    w2 = w1 - Qf @ (Qf.T @ w1)
    expected = h1.copy()
    expected[3] = np.linalg.norm(w2)
    assert np.allclose(H2[0][:, 2], expected)
    assert np.allclose(w[0], w2)
    assert np.allclose(Qf[:, 2], q2)


def test_last_norm_is_dropped_and_residual_unnormalised():
    # decomp.py:474 drops h[k] (out of range); residual is the raw w (:448-451)
    n, k = 20, 6
    A = sym_from_eigs(np.linspace(1, 2, n), 6)
    v = np.ones(n)
    Q, H, r, _ = slq.hessenberg_forward(slq.DenseOperator(A, np.float64), k, v[None], np.float64)
    assert np.allclose(np.tril(H[0], -2), 0.0)
    assert np.linalg.norm(r[0]) > 1e-3
    # the residual is orthogonal to the basis
    assert np.allclose(Q[0].T @ r[0], 0.0, atol=1e-12)


def test_k_out_of_range_raises_exceeds():
    # tests/test_decomp/test_consistency.py:50-63
    A = np.eye(13)
    for k in (-1, 14):
        with pytest.raises(ValueError, match="exceeds"):
            slq.hessenberg_forward(slq.DenseOperator(A), k, np.ones((1, 13)))


# -- integrand and estimator (funm.py:186-243, 322-335; stochtrace.py:47-50) --


def test_quadform_exact_at_full_depth():
    # tests/test_funm/test_monte_carlo_funm_sym_logdet.py:41-67
    n = 50
    A = sym_from_eigs(np.arange(1.0, 1.0 + n), 1)
    x = np.random.default_rng(1).standard_normal(n) + 10
    vals, _, _ = slq.quadform_logdet(slq.DenseOperator(A, np.float64), n - 1, x[None], np.float64)
    w, U = np.linalg.eigh(A)
    expected = x @ (U * np.log(w)) @ U.T @ x
    assert np.allclose(vals[0], expected, rtol=1e-10)


def test_logdet_estimate_matches_slogdet():
    # tests/test_funm/test_monte_carlo_funm_sym_logdet.py:16-38 (tolerance 1e-2)
    n = 200
    d = np.arange(n) / n + 1.0
    d[30:] = 0.001
    A = sym_from_eigs(d, 2)
    est, _, _, _ = slq.slq_logdet(slq.DenseOperator(A), 10, prng.prng_key(1), 10)
    expected = np.linalg.slogdet(A)[1]
    assert np.allclose(est, expected, atol=1e-2 * abs(expected), rtol=1e-2)


def test_gauss_quadrature_identity():
    # [log T]_00 = sum_j U_0j^2 log(lambda_j): dense funm vs the FP64 eigen/quadrature checker
    rng = np.random.default_rng(9)
    P, k = 4, 12
    alpha = rng.uniform(2.0, 3.0, (P, k))
    beta = rng.uniform(0.1, 0.5, (P, k - 1))
    scale = np.full(P, 3.0)
    out = slq.tridiag_eig_quadrature(alpha, beta, scale)
    for p in range(P):
        T = np.diag(alpha[p]) + np.diag(beta[p], 1) + np.diag(beta[p], -1)
        w, U = np.linalg.eigh(T)
        logT = (U * np.log(w)) @ U.T
        assert np.isclose(out[p], 9.0 * logT[0, 0], rtol=1e-12)


def test_nonpositive_ritz_value_gives_nan():
    # funm.py:201, 332: the log is unguarded
    alpha = np.array([[1.0, -2.0]])
    beta = np.array([[0.5]])
    assert np.isnan(slq.tridiag_eig_quadrature(alpha, beta, np.ones(1))[0])


def test_fp32_restatement_close_to_fp64():
    n, P, k = 256, 8, 20
    A = sym_from_eigs(np.linspace(0.5, 5.0, n), 3)
    z = prng.rademacher(prng.prng_key(4), (P, n), np.float32)
    v32, a32, b32 = slq.quadform_logdet(slq.DenseOperator(A, np.float32), k, z, np.float32)
    v64, a64, b64 = slq.quadform_logdet(slq.DenseOperator(A, np.float64), k, z, np.float64)
    assert np.allclose(v32, v64, rtol=1e-4)
    assert np.allclose(a32, a64, atol=1e-4 * np.abs(a64).max())


# -- reverse mode (decomp.py:480-600 + eigh VJP) ------------------------------


def _fd_grad(f, theta, h):
    g = []
    for i in range(len(theta)):
        e = np.zeros(len(theta))
        e[i] = h
        g.append((f(theta + e) - f(theta - e)) / (2 * h))
    return np.array(g)


def test_adjoint_matches_central_differences_rbf():
    rng = np.random.default_rng(11)
    X = rng.uniform(size=(40, 3))
    Z = prng.rademacher(prng.prng_key(3), (4, 40), np.float64)
    k = 8

    def f(theta):
        op = slq.RBFOperator(X, theta[0], theta[1], np.float64)
        vals, _, _ = slq.quadform_logdet(op, k, Z, np.float64)
        return vals.mean()

    theta = np.array([0.5, 0.1])
    _, g = slq.slq_logdet_vjp(slq.RBFOperator(X, *theta, np.float64), k, Z)
    assert np.allclose(g, _fd_grad(f, theta, 1e-5), rtol=1e-6)


def test_adjoint_dense_matches_central_differences():
    # matvec(v, A) = A v with A a parameter (test_tridiag_sym_adjoint.py style)
    n, k = 12, 5
    A0 = sym_from_eigs(np.linspace(1.0, 3.0, n), 7)
    Z = prng.rademacher(prng.prng_key(8), (2, n), np.float64)
    _, G = slq.slq_logdet_vjp(slq.DenseOperator(A0, np.float64), k, Z)
    D = np.random.default_rng(0).standard_normal((n, n))
    D = D + D.T
    h = 1e-6

    def f(A):
        vals, _, _ = slq.quadform_logdet(slq.DenseOperator(A, np.float64), k, Z, np.float64)
        return vals.mean()

    fd = (f(A0 + h * D) - f(A0 - h * D)) / (2 * h)
    assert np.isclose(np.sum(G * D), fd, rtol=1e-6)


def test_adjoint_full_depth_matches_dense_closed_form():
    # At k = n the integrand is x^T log(A(theta)) x exactly; its derivative
    # is x^T L_log(A, dA) x = sum_ij (U^T x)_i (U^T x)_j D_ij (U^T dA U)_ij
    # with D the divided differences of log at the eigenvalues of A.
    rng = np.random.default_rng(12)
    n = 10
    X = rng.uniform(size=(n, 2))
    theta = (0.7, 0.3)
    op = slq.RBFOperator(X, *theta, np.float64)
    z = prng.rademacher(prng.prng_key(5), (1, n), np.float64)
    _, g = slq.slq_logdet_vjp(op, n, z)
    A = op.dense()
    w, U = np.linalg.eigh(A)
    L = np.empty((n, n))
    for i in range(n):
        for j in range(n):
            L[i, j] = 1 / w[i] if i == j else (np.log(w[i]) - np.log(w[j])) / (w[i] - w[j])
    y = U.T @ z[0]
    d2 = np.sum((X[:, None] - X[None]) ** 2, -1)
    dA_dl = np.exp(-d2 / (2 * theta[0] ** 2)) * d2 / theta[0] ** 3
    for dA, gi in zip([dA_dl, np.eye(n)], g):
        M = U.T @ dA @ U
        assert np.isclose(np.sum(np.outer(y, y) * L * M), gi, rtol=1e-7)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_hessenberg_adjoint_random_cotangents_central_differences(seed):
    # general cotangents on every output (Q, H, r, c), the check of
    # tests/test_decomp/test_hessenberg_adjoint.py: the adjoint's (dv, dA)
    # against central differences of <cotangents, outputs> of the FP64
    # forward along a random start-vector direction and a symmetric matrix
    # direction (the adjoint applies A^T as A)
    rng = np.random.default_rng(seed)
    n, k = 12, 6
    A0 = sym_from_eigs(rng.uniform(1.0, 2.0, n), seed)
    v0 = rng.standard_normal(n)
    dQ, dH, dr, dc = (rng.standard_normal((n, k)), rng.standard_normal((k, k)),
                      rng.standard_normal(n), rng.standard_normal())
    Q, H, r, c = slq.hessenberg_forward(slq.DenseOperator(A0, np.float64), k, v0, np.float64)
    dv, dA = slq.hessenberg_adjoint(slq.DenseOperator(A0, np.float64), Q[0], H[0], r[0],
                                    float(c[0]), dQ, dH, dr, dc)

    def f(v, A):
        Q, H, r, c = slq.hessenberg_forward(slq.DenseOperator(A, np.float64), k, v, np.float64)
        return np.sum(dQ * Q[0]) + np.sum(dH * H[0]) + dr @ r[0] + dc * c[0]

    dv_dir = rng.standard_normal(n)
    D = rng.standard_normal((n, n))
    D = D + D.T
    h = 1e-6
    fd = (f(v0 + h * dv_dir, A0 + h * D) - f(v0 - h * dv_dir, A0 - h * D)) / (2 * h)
    assert np.isclose(dv @ dv_dir + np.sum(dA * D), fd, rtol=1e-6)


def test_hessenberg_adjoint_without_reorthogonalisation_agrees_in_fp64():
    # decomp.py:575: the adjoint's reorthogonalisation only stabilises; with
    # an orthonormal FP64 basis both forms give the same cotangents
    rng = np.random.default_rng(4)
    n, k = 20, 8
    A0 = sym_from_eigs(rng.uniform(1.0, 2.0, n), 4)
    v0 = rng.standard_normal(n)
    cot = (rng.standard_normal((n, k)), rng.standard_normal((k, k)), rng.standard_normal(n),
           rng.standard_normal())
    op = slq.DenseOperator(A0, np.float64)
    Q, H, r, c = slq.hessenberg_forward(op, k, v0, np.float64)
    full = slq.hessenberg_adjoint(op, Q[0], H[0], r[0], float(c[0]), *cot, reortho="full")
    none = slq.hessenberg_adjoint(op, Q[0], H[0], r[0], float(c[0]), *cot, reortho="none")
    for a, b in zip(full, none):
        assert np.allclose(a, b, rtol=1e-8, atol=1e-10)


def test_adjoint_k0_raises():
    with pytest.raises(ValueError, match="= 0"):
        slq.hessenberg_adjoint(None, np.zeros((3, 0)), np.zeros((0, 0)), np.zeros(3), 1.0,
                               np.zeros((3, 0)), np.zeros((0, 0)), np.zeros(3), 0.0)


# -- planted mistakes: the checks above must catch them ------------------------


def test_planted_mistakes_are_caught(monkeypatch):
    n = 12
    A = sym_from_eigs(np.arange(1.0, 2.0, 1 / n), 1)
    v = np.flip(np.arange(1.0, 1.0 + n))

    def relation_residual():
        k = 5
        Qt, T, r, _ = slq.tridiag_sym(slq.DenseOperator(A, np.float64), k, v[None], np.float64)
        return np.abs(A @ Qt[0].T - Qt[0].T @ T[0] - np.outer(r[0], np.eye(k)[-1])).max()

    assert relation_residual() < 1e-10
    original = slq._hessenberg_step

    # (1) wrong sign in the first projection
    def wrong_sign(Q, H, v, length, apply, idx):
        Qo, Ho, w, ln = original(Q, H, v, length, apply, idx)
        Ho[:, :, idx] *= -1.0
        return Qo, Ho, w, ln

    monkeypatch.setattr(slq, "_hessenberg_step", wrong_sign)
    assert relation_residual() > 1e-3

    # (2) norm stored one row too low (index error)
    def shifted(Q, H, v, length, apply, idx):
        Qo, Ho, w, ln = original(Q, H, v, length, apply, idx)
        if idx + 2 < Ho.shape[1]:
            Ho[:, idx + 2, idx] = Ho[:, idx + 1, idx]
            Ho[:, idx + 1, idx] = 0
        return Qo, Ho, w, ln

    monkeypatch.setattr(slq, "_hessenberg_step", shifted)
    assert relation_residual() > 1e-3


# -- funm_lanczos_sym and mean/SEM (funm.py:114-147, stochtrace.py:55-89) -------


@pytest.mark.parametrize("fun", [np.sin, np.exp, np.log])
def test_funm_lanczos_full_depth_is_exact(fun):
    # at k = n the Krylov space is everything: |v| Q^T f(T) e1 = f(A) v
    n = 12
    A = sym_from_eigs(np.linspace(0.2, 1.8, n), 3).astype(np.float64)
    v = np.random.default_rng(4).standard_normal((2, n))
    got = slq.funm_lanczos_sym(slq.DenseOperator(A, np.float64), n, v, fun, np.float64)
    w, U = np.linalg.eigh(A)
    expected = v @ (U * fun(w)) @ U.T
    assert np.allclose(got, expected, atol=1e-10)


def test_funm_lanczos_reference_configuration():
    # tests/test_funm/test_funm_lanczos_sym.py:8-36: n = 11, eigenvalues in
    # [0.01, 0.99], f = sin, 6 Lanczos steps, atol 1e-6 (float32 there)
    n = 11
    A = sym_from_eigs(np.linspace(0.01, 0.99, n), 1)
    v = np.random.default_rng(2).standard_normal((1, n)).astype(np.float32)
    got = slq.funm_lanczos_sym(slq.DenseOperator(A), 6, v, np.sin, np.float32)
    w, U = np.linalg.eigh(A.astype(np.float64))
    expected = v.astype(np.float64) @ (U * np.sin(w)) @ U.T
    assert np.allclose(got, expected, atol=1e-6)


def test_funm_lanczos_log_quadform_matches_integrand():
    # v^T (|v| Q^T log(T) e1) = |v|^2 [log T]_00: the logdet integrand
    n, k = 30, 8
    A = sym_from_eigs(np.linspace(0.5, 2.0, n), 6).astype(np.float64)
    Z = prng.rademacher(prng.prng_key(2), (3, n), np.float64)
    fv = slq.funm_lanczos_sym(slq.DenseOperator(A, np.float64), k, Z, np.log, np.float64)
    vals, _, _ = slq.quadform_logdet(slq.DenseOperator(A, np.float64), k, Z, np.float64)
    assert np.allclose(np.sum(Z * fv, axis=1), vals, rtol=1e-12)


def test_mean_and_sem_closed_form():
    s = np.array([1.0, 2.0, 4.0, 7.0])
    mean, sem = slq.estimator_mean_and_sem(s, np.float64)
    assert mean == 3.5
    assert np.isclose(sem, np.sqrt(np.mean((s - 3.5) ** 2)) / 2.0)


# -- plain Lanczos, reortho="none" (decomp.py:155-266) -------------------------


@pytest.mark.parametrize("k", [1, 5, 11])
def test_lanczos_decomposition_relation(k):
    # tests/test_decomp/test_tridiag_sym.py:43-68 with reortho="none":
    # A Q^T = Q^T T + q e_K^T, q the residual b q_k
    n = 12
    A = sym_from_eigs(np.arange(1.0, 2.0, 1 / n), 1).astype(np.float64)
    v = np.flip(np.arange(1.0, 1.0 + n))[None, :]
    Q, d, e, (qk, bk), c = slq.lanczos_forward(slq.DenseOperator(A, np.float64), k, v, np.float64)
    T = np.diag(d[0]) + np.diag(e[0], 1) + np.diag(e[0], -1)
    Qt = Q[0].T
    resid = bk[0] * qk[0]
    assert np.allclose(A @ Qt - Qt @ T - np.outer(resid, np.eye(k)[-1]), 0.0, atol=1e-10)
    assert np.isclose(c[0], 1 / np.linalg.norm(v))


def test_lanczos_matches_full_reorthogonalisation_in_fp64():
    # in exact arithmetic the three-term recurrence and Arnoldi with full
    # reorthogonalisation produce the same T; in FP64 at small depth they
    # agree to rounding
    n, k = 60, 8
    A = sym_from_eigs(np.linspace(0.5, 3.0, n), 4).astype(np.float64)
    V = np.random.default_rng(5).standard_normal((3, n))
    _, d, e, _, _ = slq.lanczos_forward(slq.DenseOperator(A, np.float64), k, V, np.float64)
    _, (fd, fe), _, _ = slq.tridiag_sym(slq.DenseOperator(A, np.float64), k,
                                        V / np.linalg.norm(V, axis=1, keepdims=True),
                                        np.float64, materialize=False)
    assert np.allclose(d, fd, atol=1e-12)
    assert np.allclose(e, fe, atol=1e-12)


def test_lanczos_full_rank_reconstruction():
    # tests/test_decomp/test_tridiag_sym.py:7-40, reortho="none" tolerance 1e-1
    n = 12
    A = sym_from_eigs(np.arange(1.0, 2.0, 1 / n), 1).astype(np.float64)
    v = np.flip(np.arange(1.0, 1.0 + n))[None, :]
    Q, d, e, _, _ = slq.lanczos_forward(slq.DenseOperator(A, np.float64), n, v, np.float64)
    T = np.diag(d[0]) + np.diag(e[0], 1) + np.diag(e[0], -1)
    assert np.allclose(Q[0].T @ T @ Q[0], A, atol=1e-1, rtol=1e-1)


# -- plain Lanczos adjoint (decomp.py:268-348) ----------------------------------


def _lanczos_outputs(A, v, k):
    Q, d, e, (qk, bk), _ = slq.lanczos_forward(slq.DenseOperator(A, np.float64), k, v[None],
                                              np.float64)
    xs = np.concatenate([Q[0], qk])
    betas = np.concatenate([e[0], bk])
    return xs, d[0], betas


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_lanczos_adjoint_vjp_random_cotangents(seed):
    # tests/test_decomp/test_tridiag_sym_adjoint.py:7-50 (reortho="none", n = 10,
    # k = 4): the custom VJP against differentiation of the forward pass, here
    # central differences of the FP64 restatement along random directions of
    # (vector, symmetric matrix), for random cotangents on every output
    n, k = 10, 4
    rng = np.random.default_rng(seed)
    A0 = sym_from_eigs(rng.uniform(size=n) + 1.0, seed).astype(np.float64)
    v0 = rng.standard_normal(n)
    dxs = rng.standard_normal((k + 1, n))
    da = rng.standard_normal(k)
    db = rng.standard_normal(k)

    def f(v, A):
        xs, d, betas = _lanczos_outputs(A, v, k)
        return np.sum(dxs * xs) + da @ d + db @ betas

    xs, d, betas = _lanczos_outputs(A0, v0, k)
    gv, gA, _ = slq.lanczos_adjoint(slq.DenseOperator(A0, np.float64), d, betas, xs, da, db, dxs,
                                    np.linalg.norm(v0))
    dv = rng.standard_normal(n)
    D = rng.standard_normal((n, n))
    D = D + D.T
    h = 1e-6
    fd = (f(v0 + h * dv, A0 + h * D) - f(v0 - h * dv, A0 - h * D)) / (2 * h)
    assert np.isclose(gv @ dv + np.sum(gA * D), fd, rtol=1e-6), (gv @ dv + np.sum(gA * D), fd)


@pytest.mark.parametrize("k", [4, 8])
def test_lanczos_adjoint_matches_arnoldi_adjoint(k):
    # both reverse passes differentiate the same T(theta) in exact arithmetic;
    # in FP64 at depths where Lanczos is still orthogonal they agree to rounding
    X = np.random.default_rng(1).uniform(size=(200, 3))
    Z = prng.rademacher(prng.prng_key(3), (3, 200), np.float64)
    op = slq.RBFOperator(X, 0.5, 0.1, np.float64)
    e1, g1 = slq.slq_logdet_lanczos_vjp(op, k, Z)
    e2, g2 = slq.slq_logdet_vjp(op, k, Z)
    assert np.isclose(e1, e2, rtol=1e-12)
    assert np.allclose(g1, g2, rtol=1e-10)


def test_lanczos_adjoint_rbf_central_differences():
    # depth 16 on 40 points: past the loss of orthogonality, so the Lanczos
    # value differs from Arnoldi's and only its own derivative matches
    rng = np.random.default_rng(11)
    X = rng.uniform(size=(40, 3))
    Z = prng.rademacher(prng.prng_key(3), (4, 40), np.float64)
    k = 16

    def f(theta):
        op = slq.RBFOperator(X, theta[0], theta[1], np.float64)
        vals, _, _ = slq.quadform_logdet_lanczos(op, k, Z, np.float64)
        return vals.mean()

    theta = np.array([0.5, 0.1])
    _, g = slq.slq_logdet_lanczos_vjp(slq.RBFOperator(X, *theta, np.float64), k, Z)
    assert np.allclose(g, _fd_grad(f, theta, 1e-5), rtol=1e-5)


def test_lanczos_adjoint_planted_mistakes_are_caught(monkeypatch):
    n, k = 10, 4
    rng = np.random.default_rng(1)
    A0 = sym_from_eigs(rng.uniform(size=n) + 1.0, 1).astype(np.float64)
    v0 = rng.standard_normal(n)
    dxs = rng.standard_normal((k + 1, n))
    da = rng.standard_normal(k)
    db = rng.standard_normal(k)
    dv = rng.standard_normal(n)
    D = rng.standard_normal((n, n))
    D = D + D.T

    def mismatch():
        xs, d, betas = _lanczos_outputs(A0, v0, k)
        gv, gA, _ = slq.lanczos_adjoint(slq.DenseOperator(A0, np.float64), d, betas, xs, da, db,
                                        dxs, np.linalg.norm(v0))
        h = 1e-6

        def f(v, A):
            xs, d, betas = _lanczos_outputs(A, v, k)
            return np.sum(dxs * xs) + da @ d + db @ betas

        fd = (f(v0 + h * dv, A0 + h * D) - f(v0 - h * dv, A0 - h * D)) / (2 * h)
        return abs(gv @ dv + np.sum(gA * D) - fd) / abs(fd)

    assert mismatch() < 1e-6
    original = slq._lanczos_adjoint_step

    # (1) wrong sign of the b * lambda_plus term
    def wrong_sign(operator, xi, lam_plus, x, xplus, a, b, da, db, dx):
        return original(operator, xi, -lam_plus, x, xplus, a, b, da, db, dx)

    monkeypatch.setattr(slq, "_lanczos_adjoint_step", wrong_sign)
    assert mismatch() > 1e-3

    # (2) neighbouring vectors swapped (index error)
    def swapped(operator, xi, lam_plus, x, xplus, a, b, da, db, dx):
        return original(operator, xi, lam_plus, xplus, x, a, b, da, db, dx)

    monkeypatch.setattr(slq, "_lanczos_adjoint_step", swapped)
    assert mismatch() > 1e-3
