"""Pinning of the low-rank / CG / log-density restatement (oracle/lowrank.py).

The reference's own checks (tests/test_low_rank/test_cholesky.py,
tests/test_low_rank/test_preconditioner.py) are restated on matrices of the
same construction (random orthogonal Q times prescribed eigenvalues; JAX's
random streams are not reproducible here, so the matrices differ), plus what
the mathematics fixes: exact reconstruction at full rank, agreement with
``numpy.linalg.cholesky`` without pivoting, the Woodbury identity against a
dense solve, CG against a dense solve and its n-step termination, the Gaussian
log-density against ``scipy.stats``, and planted mistakes.
"""

import numpy as np
import pytest
import scipy.stats

from oracle import lowrank as olr


def sym_from_eigs(eigvals, seed):
    n = len(eigvals)
    Q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((n, n)))
    return (Q * np.asarray(eigvals)) @ Q.T


def _pivot(A, rank, dtype=np.float64):
    L, info, _ = olr.cholesky_partial_pivot(A, rank, dtype)
    return L, info


@pytest.mark.parametrize("chol", [olr.cholesky_partial, _pivot])
def test_full_rank_cholesky_reconstructs_matrix(chol):
    # test_cholesky.py:16-28 (n = 5, eigenvalues 1 + U(0, 1)), in FP64
    n = 5
    cov = sym_from_eigs(1.0 + np.random.default_rng(2).uniform(size=n), 2)
    L, _ = chol(cov, n, np.float64)
    tol = np.finfo(np.float64).eps * 16
    assert np.allclose(L @ L.T, cov, atol=tol, rtol=tol)


@pytest.mark.parametrize("chol", [olr.cholesky_partial, _pivot])
def test_output_the_right_shapes(chol):
    # test_cholesky.py:31-42
    n, rank = 4, 4
    cov = sym_from_eigs(0.1 + np.random.default_rng(1).uniform(size=n), 1)
    L, _ = chol(cov, rank, np.float64)
    assert L.shape == (n, rank)


def test_full_rank_nopivot_matches_cholesky():
    # test_cholesky.py:45-62
    n = 10
    cov = sym_from_eigs(0.01 + np.random.default_rng(2).uniform(size=n), 2)
    reference = np.linalg.cholesky(cov)
    received, _ = _pivot(cov, n)
    assert not np.allclose(received, reference)
    received, _ = olr.cholesky_partial(cov, n, np.float64)
    assert np.allclose(received, reference, atol=1e-6)


def test_pivoting_improves_the_estimate():
    # test_cholesky.py:65-81
    n, rank = 10, 5
    cov = sym_from_eigs(0.1 + np.random.default_rng(1).uniform(size=n), 1)
    nopivot, _ = olr.cholesky_partial(cov, rank, np.float64)
    pivot, _ = _pivot(cov, rank)
    assert np.linalg.norm(cov - pivot @ pivot.T) < np.linalg.norm(cov - nopivot @ nopivot.T)


def test_pivot_order_and_success_flag():
    # the first pivot is the largest diagonal entry; the permutation is one;
    # a rank-deficient matrix (rank 3) fails at step 4 (l_ii^2 <= 0)
    n = 8
    rng = np.random.default_rng(3)
    B = rng.standard_normal((n, 3))
    A = B @ B.T
    L, info, P = olr.cholesky_partial_pivot(A, 3, np.float64)
    assert P[0] == int(np.argmax(np.diag(A)))
    assert sorted(P.tolist()) == list(range(n))
    assert info["success"] and np.allclose(L @ L.T, A, atol=1e-10)
    _, info4, _ = olr.cholesky_partial_pivot(A + 0.0, 5, np.float64)
    assert not info4["success"]


@pytest.mark.parametrize("rank", [1, 11])
def test_rank_out_of_range_raises(rank):
    A = np.eye(10)
    for f in (olr.cholesky_partial, olr.cholesky_partial_pivot):
        with pytest.raises(ValueError, match="Rank"):
            f(A, rank - 1 if rank == 1 else rank)


def test_preconditioner_solves_correctly():
    # test_preconditioner.py:7-40: full-rank factor of an ill-conditioned
    # matrix (eigenvalues 1.5^k, k = -5..4), shift 0.1, rtol 1e-2 -- and, in
    # FP64, the Woodbury identity to rounding
    n = 10
    cov = sym_from_eigs(1.5 ** np.arange(-n // 2, n // 2, step=1.0), 1)
    L, _ = olr.cholesky_partial(cov, n, np.float64)
    assert np.allclose(L @ L.T, cov)
    b = np.arange(1.0, 1 + n)
    b /= np.linalg.norm(b)
    expected = np.linalg.solve(cov + 0.1 * np.eye(n), b)
    received = olr.preconditioner_solve(L, b, 0.1, np.float64)
    assert np.allclose(received, expected, rtol=1e-2)
    assert np.allclose(received, expected, rtol=1e-10, atol=1e-12)
    # low rank: (s I + L L^T)^{-1} exactly, whatever L is
    Lr = L[:, :4]
    exp_r = np.linalg.solve(0.1 * np.eye(n) + Lr @ Lr.T, b)
    assert np.allclose(olr.preconditioner_solve(Lr, b, 0.1, np.float64), exp_r, rtol=1e-10)


def test_cg_solves_spd_systems():
    n, P = 60, 3
    A = sym_from_eigs(np.linspace(0.5, 20.0, n), 4)
    B = np.random.default_rng(5).standard_normal((P, n))
    X, k = olr.cg(lambda V: V @ A, B, tol=1e-10, dtype=np.float64)
    assert np.allclose(X, np.linalg.solve(A, B.T).T, rtol=1e-8, atol=1e-10)
    # residual criterion: |r| <= tol |b| for every column
    R = B - X @ A
    assert np.all(np.linalg.norm(R, axis=1) <= 1e-10 * np.linalg.norm(B, axis=1) * 1.0001)
    assert np.all(k <= n + 5)


def test_cg_terminates_in_at_most_n_steps_on_few_distinct_eigenvalues():
    # CG converges in at most d steps for d distinct eigenvalues (exact arithmetic)
    n = 40
    A = sym_from_eigs(np.repeat([1.0, 3.0, 7.0], [10, 15, 15]), 6)
    b = np.random.default_rng(1).standard_normal((1, n))
    X, k = olr.cg(lambda V: V @ A, b, tol=1e-12, dtype=np.float64)
    assert k[0] <= 4
    assert np.allclose(X[0], np.linalg.solve(A, b[0]), rtol=1e-9)


def test_cg_freezes_converged_columns_and_honours_maxiter():
    n = 30
    A = sym_from_eigs(np.linspace(1.0, 100.0, n), 7)
    B = np.stack([np.eye(n)[0] * 0.0, np.random.default_rng(2).standard_normal(n)])
    X, k = olr.cg(lambda V: V @ A, B, tol=1e-8, dtype=np.float64)
    assert k[0] == 0 and np.all(X[0] == 0.0)  # b = 0: rs = 0 <= atol2 at once
    assert k[1] > 0
    _, k3 = olr.cg(lambda V: V @ A, B, tol=1e-12, maxiter=3, dtype=np.float64)
    assert k3[1] == 3


def test_preconditioned_cg_converges_faster():
    # a kernel matrix plus a small nugget: the pivoted partial Cholesky of the
    # kernel is an effective preconditioner (low_rank.py:7-53)
    rng = np.random.default_rng(4)
    x = rng.uniform(size=(200, 2))
    d2 = np.sum((x[:, None] - x[None]) ** 2, -1)
    K = np.exp(-d2 / (2 * 0.3**2))
    s = 1e-2
    A = K + s * np.eye(200)
    b = rng.standard_normal((1, 200))
    L, _, _ = olr.cholesky_partial_pivot(K, 40, np.float64)
    X0, k0 = olr.cg(lambda V: V @ A, b, tol=1e-8, dtype=np.float64)
    X1, k1 = olr.cg(lambda V: V @ A, b, tol=1e-8, dtype=np.float64,
                    M=lambda R: np.stack([olr.preconditioner_solve(L, r, s, np.float64) for r in R]))
    exact = np.linalg.solve(A, b[0])
    assert np.allclose(X1[0], exact, rtol=1e-5)
    assert np.allclose(X0[0], exact, rtol=1e-5)
    assert k1[0] < k0[0] / 2


def test_gaussian_logpdf_matches_scipy():
    # tutorials/8_gaussian_logpdf.py:55-66: the dense reference
    n = 30
    xs = np.linspace(0.0, 1.0, n)
    K = np.exp(-((xs[:, None] - xs[None]) ** 2) / 2.0) + np.exp(-2.0) * np.eye(n)
    y = np.random.default_rng(1).standard_normal(n)
    logdet = np.linalg.slogdet(K)[1]
    got = olr.gaussian_logpdf(y, logdet, lambda v: np.linalg.solve(K, v))
    ref = scipy.stats.multivariate_normal.logpdf(y, mean=np.zeros(n), cov=K)
    assert np.isclose(got, ref, rtol=1e-10)


def test_planted_mistakes_are_caught(monkeypatch):
    n = 6
    cov = sym_from_eigs(1.0 + np.random.default_rng(2).uniform(size=n), 2)

    def err():
        L, _, _ = olr.cholesky_partial_pivot(cov, n, np.float64)
        return np.abs(L @ L.T - cov).max()

    assert err() < 1e-12
    # (1) rows of L not swapped with the permutation
    original = olr._swap
    calls = {"n": 0}

    def swap_only_perm(a, i, j):
        calls["n"] += 1
        if a.ndim == 1:
            original(a, i, j)

    monkeypatch.setattr(olr, "_swap", swap_only_perm)
    assert err() > 1e-3
    monkeypatch.setattr(olr, "_swap", original)
    # (2) preconditioner with the shift applied once instead of twice
    L, _ = olr.cholesky_partial(cov, n, np.float64)
    b = np.ones(n)
    good = olr.preconditioner_solve(L, b, 0.5, np.float64)
    assert np.allclose(good, np.linalg.solve(cov + 0.5 * np.eye(n), b))
    wrong = b - L @ np.linalg.solve(np.eye(n) + L.T @ L / 0.5, L.T @ b / 0.5) / 0.5
    assert not np.allclose(wrong, np.linalg.solve(cov + 0.5 * np.eye(n), b), rtol=1e-3)
