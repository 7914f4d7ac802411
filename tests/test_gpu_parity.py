"""GPU parity of the HIP path against the oracle (run on an MI355X).

Tolerances (float32 path vs the float32 restatement of the same formulas; the
two differ only in summation order):
  * probes: bit-exact;
  * block matvec: normwise relative error <= 2e-6 vs FP64;
  * alpha/beta: |delta| <= 1e-5 * max|alpha|;
  * per-probe and mean log-determinant: relative 1e-5.
"""

import numpy as np
import pytest
import torch

from matfree_amd import decomp, funm, operators, problems, prng, stochtrace
from oracle import prng as oprng
from oracle import slq as oslq

pytestmark = pytest.mark.gpu


def sym_from_eigs(eigvals, seed):
    n = len(eigvals)
    Q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((n, n)))
    return ((Q * eigvals) @ Q.T).astype(np.float32)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ---------------------------------------------------------------------------
# a1: probes
# ---------------------------------------------------------------------------


@pytest.mark.parametrize(
    "seed,num,n", [(0, 64, 1024), (1, 3, 1001), (1, 10000, 2), (2**32 + 17, 5, 7), (123, 1, 1)]
)
def test_probes_bit_exact(cuda, seed, num, n):
    key = prng.PRNGKey(seed)
    z = stochtrace.sampler_signs(np.zeros(n, np.float32), num=num).flat(key, cuda)
    expected = oprng.rademacher(oprng.prng_key(seed), (num, n), np.float32)
    np.testing.assert_array_equal(z.cpu().numpy(), expected)


def test_probes_readme_known_answer(cuda):
    z = stochtrace.sampler_signs(np.zeros(2, np.float32), num=10_000)(prng.PRNGKey(1))
    z = z.cpu().numpy()
    A = np.arange(12.0, dtype=np.float32).reshape(6, 2)
    vals = np.einsum("pi,ij,pj->p", z, A.T @ A, z).astype(np.float32)
    assert int(np.sum(z[:, 0] == z[:, 1])) == 4980
    assert float(np.mean(vals, dtype=np.float32)) == 504.0


def test_probe_pytree_unflatten(cuda):
    s = stochtrace.sampler_signs({"a": np.ones(3), "b": np.ones((2, 2))}, num=4)
    out = s(prng.PRNGKey(3))
    flat = oprng.rademacher(oprng.prng_key(3), (4, 7), np.float32)
    np.testing.assert_array_equal(out["a"].cpu().numpy(), flat[:, :3])
    np.testing.assert_array_equal(out["b"].cpu().numpy(), flat[:, 3:].reshape(4, 2, 2))


# ---------------------------------------------------------------------------
# a9: block matvec
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("n,P", [(1024, 64), (100, 3), (12, 1), (300, 130)])
def test_dense_matvec(cuda, n, P):
    rng = np.random.default_rng(n + P)
    A = rng.standard_normal((n, n)).astype(np.float32)
    V = rng.standard_normal((P, n)).astype(np.float32)
    W = operators.dense_matvec(torch.as_tensor(V, device=cuda), torch.as_tensor(A, device=cuda))
    expected = V.astype(np.float64) @ A.astype(np.float64).T
    assert rel(W.cpu().numpy(), expected) < 2e-6


@pytest.mark.parametrize("n,P,dim", [(1000, 16, 3), (257, 5, 1), (64, 64, 4)])
def test_rbf_matvec(cuda, n, P, dim):
    X = problems.rbf_points(n, seed=n, dim=dim)
    V = np.random.default_rng(1).standard_normal((P, n)).astype(np.float32)
    op = operators.RBF(X)
    W = op(torch.as_tensor(V, device=cuda), 0.5, 1e-2)
    K = problems.rbf_dense64(X, 0.5, 1e-2)
    assert rel(W.cpu().numpy(), V.astype(np.float64) @ K) < 2e-6


@pytest.mark.parametrize("precision", ["fp32", "bf16x2"])
@pytest.mark.parametrize("n,P,dim", [(3000, 1, 3), (5000, 2, 1), (777, 3, 2), (20000, 4, 3),
                                     (50, 1, 3)])
def test_rbf_matvec_few_vectors(cuda, precision, n, P, dim):
    # up to four vectors (CG's right-hand sides) run the FP32 row kernel,
    # whatever the requested precision; several column slices at n >= 2048
    X = problems.rbf_points(n, seed=n, dim=dim)
    V = np.random.default_rng(3).standard_normal((P, n)).astype(np.float32)
    W = operators.RBF(X, precision=precision)(torch.as_tensor(V, device=cuda), 0.4, 0.1)
    K = problems.rbf_dense64(X, 0.4, 0.1)
    assert rel(W.cpu().numpy(), V.astype(np.float64) @ K) < 2e-6


def test_rbf_matvec_few_vectors_dim4(cuda):
    X = problems.rbf_points(2500, seed=1, dim=4)
    V = np.random.default_rng(4).standard_normal((2, 2500)).astype(np.float32)
    W = operators.RBF(X)(torch.as_tensor(V, device=cuda), 0.6, 0.05)
    K = problems.rbf_dense64(X, 0.6, 0.05)
    assert rel(W.cpu().numpy(), V.astype(np.float64) @ K) < 2e-6


@pytest.mark.parametrize("precision,P,dim,tol", [
    ("fp32", 8, 3, 2e-6), ("fp32", 2, 3, 2e-6), ("fp32", 2, 4, 2e-6), ("fp32", 40, 2, 2e-6),
    ("bf16x2", 16, 3, 4e-3), ("bf16x2", 128, 3, 4e-3), ("bf16x3", 64, 3, 3e-5)])
def test_rbf_matvec_far_from_origin(cuda, precision, P, dim, tol):
    # points in [100, 101]^dim: every kernel takes them relative to point 0,
    # so they keep the accuracy of the unit cube (scaled coordinates of ~170
    # would lose ~1e-5 of each difference in FP32, and the BF16 kernels'
    # expanded exponents ~1e-2)
    n = 1500
    X = (problems.rbf_points(n, seed=9, dim=dim) + np.float32(100.0)).astype(np.float32)
    V = np.random.default_rng(5).standard_normal((P, n)).astype(np.float32)
    W = operators.RBF(X, precision=precision)(torch.as_tensor(V, device=cuda), 0.5, 1e-2)
    K = problems.rbf_dense64(X, 0.5, 1e-2)
    assert rel(W.cpu().numpy(), V.astype(np.float64) @ K) < tol


@pytest.mark.parametrize("precision,tol", [("bf16x3", 3e-5), ("bf16x2", 4e-3), ("bf16", 6e-3)])
@pytest.mark.parametrize("n,P,dim", [(1000, 16, 3), (256, 128, 3), (130, 70, 3), (333, 40, 1),
                                     (96, 200, 2)])
def test_rbf_matvec_bf16(cuda, precision, tol, n, P, dim):
    # split hi+lo BF16 recovers ~16 significant bits; plain BF16 keeps 8
    X = problems.rbf_points(n, seed=3, dim=dim)
    V = np.random.default_rng(2).standard_normal((P, n)).astype(np.float32)
    W = operators.RBF(X, precision=precision)(torch.as_tensor(V, device=cuda), 0.5, 1e-2)
    K = problems.rbf_dense64(X, 0.5, 1e-2)
    assert rel(W.cpu().numpy(), V.astype(np.float64) @ K) < tol


@pytest.mark.parametrize("precision,tol", [("bf16x3", 3e-5), ("bf16", 6e-3)])
def test_dense_matvec_bf16(cuda, precision, tol):
    n, P = 520, 40
    A = problems.dense_spd(n, seed=1, shift=1.0)
    V = np.random.default_rng(4).standard_normal((P, n)).astype(np.float32)
    op = operators.DenseMatvec(precision=precision)
    W = op(torch.as_tensor(V, device=cuda), torch.as_tensor(A, device=cuda))
    assert rel(W.cpu().numpy(), V.astype(np.float64) @ A.astype(np.float64).T) < tol


def test_single_vector_matvec(cuda):
    A = sym_from_eigs(np.linspace(1, 2, 33), 0)
    v = np.arange(33, dtype=np.float32)
    w = operators.dense_matvec(torch.as_tensor(v, device=cuda), torch.as_tensor(A, device=cuda))
    assert w.shape == (33,)
    assert rel(w.cpu().numpy(), A.astype(np.float64) @ v) < 2e-6


# ---------------------------------------------------------------------------
# a6-a8: Arnoldi / tridiagonalisation
# ---------------------------------------------------------------------------


def test_full_rank_reconstruction_is_exact(cuda):
    # tests/test_decomp/test_tridiag_sym.py:7-38
    n = 12
    A = sym_from_eigs(np.arange(1.0, 2.0, 1 / n), 1)
    vector = np.flip(np.arange(1.0, 1.0 + n)).astype(np.float32).copy()
    Q, T, _, ln = decomp.tridiag_sym(n, materialize=True)(
        operators.dense_matvec, torch.as_tensor(vector, device=cuda), torch.as_tensor(A, device=cuda)
    )
    Q = Q.cpu().numpy().astype(np.float64)
    T = T.cpu().numpy().astype(np.float64)
    assert np.allclose(Q.T @ T @ Q, A, atol=1e-5, rtol=1e-5)
    assert np.allclose(Q.T @ Q, np.eye(n), atol=1e-5)
    assert np.allclose(Q @ Q.T, np.eye(n), atol=1e-5)
    assert np.isclose(float(ln), np.linalg.norm(vector), rtol=1e-6)


@pytest.mark.parametrize("k", [1, 5, 11])
def test_mid_rank_lanczos_relation(cuda, k):
    # tests/test_decomp/test_tridiag_sym.py:43-65
    n = 12
    A = sym_from_eigs(np.arange(1.0, 2.0, 1 / n), 1)
    vector = np.flip(np.arange(1.0, 1.0 + n)).astype(np.float32).copy()
    Q, T, q, _ = decomp.tridiag_sym(k)(operators.dense_matvec, vector, A)
    Q, T, q = (x.cpu().numpy().astype(np.float64) for x in (Q, T, q))
    tol = 10 * np.sqrt(np.finfo(np.float32).eps)
    assert np.allclose(A @ Q.T - Q.T @ T - np.outer(q, np.eye(k)[-1]), 0.0, atol=tol)


@pytest.mark.parametrize("k", [0, 5, 9])
def test_hessenberg_decomposition_nonsymmetric(cuda, k):
    # tests/test_decomp/test_hessenberg.py:7-38 (general square A)
    n = 10
    rng = np.random.default_rng(1)
    A = rng.standard_normal((n, n)).astype(np.float32)
    v = rng.standard_normal(n).astype(np.float32)
    Q, H, r, c = decomp.hessenberg(k, reortho="full")(operators.dense_matvec, v, A)
    assert Q.shape == (k, n) and H.shape == (k, k) and r.shape == (n,) and c.shape == ()
    Q, H, r = (x.cpu().numpy().astype(np.float64) for x in (Q, H, r))
    tol = 10 * np.sqrt(np.finfo(np.float32).eps)
    if k:
        ek = np.eye(k)[-1]
        assert np.allclose(A @ Q.T - Q.T @ H - np.outer(r, ek), 0.0, atol=tol)
        assert np.allclose(Q @ Q.T, np.eye(k), atol=tol)
        assert np.allclose(Q.T[:, 0], float(c) * v, atol=tol)
    else:
        assert np.allclose(r, v)
        assert np.isclose(float(c), 1 / np.linalg.norm(v), rtol=1e-6)


def fp32_spread(A, V0, k, ref_Q, ref_H, perms=3):
    """Largest distance of the FP32 restatement from FP64 over element orders.

    Reordering the elements (a symmetric permutation of A and the start
    vectors) changes only the FP32 summation order, so this measures how far
    any correct FP32 implementation of the reference formulas may land from
    the FP64 result.
    """
    n = A.shape[0]
    worst_q, worst_h = 0.0, 0.0
    for seed in range(perms + 1):
        perm = np.arange(n) if seed == 0 else np.random.default_rng(seed).permutation(n)
        inv = np.argsort(perm)
        Q, H, _, _ = oslq.hessenberg_forward(oslq.DenseOperator(A[perm][:, perm]), k,
                                             V0[:, perm], np.float32)
        worst_q = max(worst_q, rel(Q[:, inv, :], ref_Q))
        worst_h = max(worst_h, float(np.abs(H - ref_H).max()))
    return worst_q, worst_h


@pytest.mark.parametrize("n,P,k", [(1024, 64, 30), (1001, 7, 20), (1002, 9, 24), (200, 3, 70),
                                   (64, 5, 48)])
def test_hessenberg_matches_oracle(cuda, n, P, k):
    # n = 1024: aligned float4 streaming; 1002 and 1001: float4 groups from each row's
    # 16-byte boundary for the dots and projection, float2 / single floats elsewhere
    A = problems.dense_spd(n, seed=0, shift=1.0)
    V = oprng.rademacher(oprng.prng_key(n), (P, n), np.float32)
    V0 = V / np.sqrt(n, dtype=np.float32)
    Q, H, r, c = decomp.arnoldi(operators.dense_matvec, torch.as_tensor(V0, device=cuda), k,
                                (torch.as_tensor(A, device=cuda),))
    oQ, oH, orr, oc = oslq.hessenberg_forward(oslq.DenseOperator(A), k, V0, np.float32)
    dQ, dH, _, _ = oslq.hessenberg_forward(oslq.DenseOperator(A, np.float64), k, V0, np.float64)
    H = H.cpu().numpy()
    Qg = Q.permute(1, 2, 0).cpu().numpy()
    # H (alpha, beta and the super-diagonal) against the FP32 restatement:
    # within 1e-5 of its largest entry, or within the FP32 order spread.
    spread_q, spread_h = fp32_spread(A, V0, k, dQ, dH)
    h_err = float(np.abs(H - dH).max())
    assert h_err <= max(1e-5 * np.abs(dH).max(), 1.5 * spread_h), (h_err, spread_h)
    q_err = rel(Qg, dQ)
    assert q_err <= 1.5 * spread_q + 1e-6, (q_err, spread_q)
    assert np.allclose(c.cpu().numpy(), oc, rtol=1e-6)
    # invariants at this size
    Qd = Qg.astype(np.float64)
    for p in range(min(P, 4)):
        G = Qd[p].T @ Qd[p]
        assert np.abs(G - np.eye(k)).max() < 1e-5
        e_k = np.eye(k)[-1]
        resid = A.astype(np.float64) @ Qd[p] - Qd[p] @ H[p] - np.outer(r[p].cpu().numpy(), e_k)
        assert np.abs(resid).max() < 10 * np.sqrt(np.finfo(np.float32).eps)


# ---------------------------------------------------------------------------
# a2-a5: integrand, quadrature and estimator
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("precision,tol_full,tol_none",
                         [("bf16x3", 1e-4, 3e-3), ("bf16x2", 1e-3, 0.1), ("bf16", 0.2, 0.2)])
@pytest.mark.parametrize("reortho", ["full", "none"])
def test_bf16_rbf_logdet_tracks_fp32(cuda, precision, tol_full, tol_none, reortho):
    # multi-step paths in BF16: after the first matvec the normalisation
    # writes the next Krylov vector's BF16 planes itself (40 probes pad to 64,
    # n = 4000 pads to whole 32-column tiles).  Plain BF16 rounds K itself to
    # 8 bits, which moves the small eigenvalues at s2 = 0.1 by tens of percent
    # (measured: 9% per probe); plain Lanczos in FP32 moves by ~3e-4 per probe
    # under any change of rounding at this depth (measured with bf16x3), and by
    # ~4% when every Krylov vector is rounded to BF16 (bf16x2: it loses
    # orthogonality sooner); a misplaced plane would be off by far more.  The
    # fused planes are bit-identical to the split pass (same log-determinant
    # to the last bit at n = 32768, 128 probes, both forms).
    tol = tol_full if reortho == "full" else tol_none
    n, P, k = 4000, 40, 24
    X = problems.rbf_points(n, seed=5)
    Z = torch.as_tensor(oprng.rademacher(oprng.prng_key(3), (P, n), np.float32), device=cuda)
    ref, _ = funm.logdet_per_probe(operators.RBF(X), Z, k, (0.5, 0.1), reortho=reortho)
    got, _ = funm.logdet_per_probe(operators.RBF(X, precision=precision), Z, k, (0.5, 0.1),
                                   reortho=reortho)
    ref, got = ref.cpu().numpy(), got.cpu().numpy()
    assert np.max(np.abs(got - ref) / np.abs(ref)) <= tol


@pytest.mark.parametrize("precision,tol", [("bf16x3", 1e-4), ("bf16", 3e-2)])
def test_bf16_dense_logdet_tracks_fp32(cuda, precision, tol):
    n, P, k = 2000, 24, 20
    A = torch.as_tensor(problems.dense_spd(n, seed=2, shift=1.0), device=cuda)
    Z = torch.as_tensor(oprng.rademacher(oprng.prng_key(4), (P, n), np.float32), device=cuda)
    ref, _ = funm.logdet_per_probe(operators.DenseMatvec(), Z, k, (A,))
    got, _ = funm.logdet_per_probe(operators.DenseMatvec(precision=precision), Z, k, (A,))
    ref, got = ref.cpu().numpy(), got.cpu().numpy()
    assert np.max(np.abs(got - ref) / np.abs(ref)) <= tol


def test_logdet_c1_matches_oracle_and_cholesky(cuda):
    # Plain SLQ logdet: A = B B^T / n + I, n = 1024, 64 probes, depth 30
    n, P, k = 1024, 64, 30
    A = problems.dense_spd(n, seed=0, shift=1.0)
    key = prng.PRNGKey(0)
    Z = stochtrace.sampler_signs(np.zeros(n, np.float32), num=P).flat(key, cuda)
    per, mean, alpha, beta = funm.logdet_per_probe(
        operators.dense_matvec, Z, k, (torch.as_tensor(A, device=cuda),), want_coeffs=True
    )
    o_est, o_vals, o_alpha, o_beta = oslq.slq_logdet(oslq.DenseOperator(A), k, oprng.prng_key(0), P)
    scale = np.abs(o_alpha).max()
    assert np.abs(alpha.cpu().numpy() - o_alpha).max() <= 1e-5 * scale
    assert np.abs(beta.cpu().numpy() - o_beta).max() <= 1e-5 * scale
    assert np.allclose(per.cpu().numpy(), o_vals, rtol=1e-5)
    assert abs(float(mean) - float(o_est)) <= 1e-5 * abs(float(o_est))
    # FP64 eigen/quadrature of the GPU's own tridiagonal matches numpy's eigh
    q = oslq.tridiag_eig_quadrature(alpha.cpu().numpy(), beta.cpu().numpy(),
                                    np.sqrt(np.full(P, n, np.float32)))
    assert np.allclose(per.cpu().numpy(), q, rtol=1e-10)
    chol = oslq.logdet_cholesky(A)
    assert abs(float(mean) - chol) / abs(chol) < 1e-2


def test_estimator_api_end_to_end(cuda):
    # tutorials/1_log_determinants.py:29-35 composition
    n, P, k = 512, 32, 20
    A = problems.dense_spd(n, seed=3, shift=1.0)
    tridiag = decomp.tridiag_sym(k)
    integrand = funm.monte_carlo_funm_sym_logdet(tridiag)
    sampler = stochtrace.sampler_signs(np.zeros(n, np.float32), num=P)
    estimate = stochtrace.estimator_monte_carlo(integrand, sampler)
    got = float(estimate(operators.dense_matvec, prng.PRNGKey(5), torch.as_tensor(A, device=cuda)))
    o_est, _, _, _ = oslq.slq_logdet(oslq.DenseOperator(A), k, oprng.prng_key(5), P)
    assert abs(got - float(o_est)) <= 1e-5 * abs(float(o_est))


def test_logdet_exact_for_full_depth(cuda):
    # tests/test_funm/test_monte_carlo_funm_sym_logdet.py:41-67
    n = 50
    A = sym_from_eigs(np.arange(1.0, 1.0 + n), 1)
    x = (np.random.default_rng(1).standard_normal(n) + 10).astype(np.float32)
    integrand = funm.monte_carlo_funm_sym_logdet(decomp.tridiag_sym(n - 1))
    got = float(integrand(operators.dense_matvec, x, A))
    w, U = np.linalg.eigh(A.astype(np.float64))
    expected = x @ (U * np.log(w)) @ U.T @ x
    assert np.isclose(got, expected, rtol=2e-5)


def test_logdet_spd_matches_slogdet(cuda):
    # tests/test_funm/test_monte_carlo_funm_sym_logdet.py:16-38 (Rademacher probes here)
    n = 200
    d = np.arange(n) / n + 1.0
    d[30:] = 0.001
    A = sym_from_eigs(d, 2)
    integrand = funm.monte_carlo_funm_sym_logdet(decomp.tridiag_sym(10))
    est = stochtrace.estimator_monte_carlo(integrand, stochtrace.sampler_signs(np.ones(n), num=10))
    got = float(est(operators.dense_matvec, prng.PRNGKey(1), A))
    expected = np.linalg.slogdet(A.astype(np.float64))[1]
    assert np.allclose(got, expected, atol=1e-2 * abs(expected), rtol=1e-2)


def test_rbf_logdet_matches_oracle(cuda):
    n, P, k = 512, 16, 25
    X = problems.rbf_points(n, seed=1)
    Z = oprng.rademacher(oprng.prng_key(9), (P, n), np.float32)
    per, mean = funm.logdet_per_probe(operators.RBF(X), torch.as_tensor(Z, device=cuda), k,
                                      (0.5, 0.1))
    o_vals, _, _ = oslq.quadform_logdet(oslq.RBFOperator(X, 0.5, 0.1, np.float32), k, Z, np.float32)
    assert np.allclose(per.cpu().numpy(), o_vals, rtol=1e-4)
    o64, _, _ = oslq.quadform_logdet(oslq.RBFOperator(X, 0.5, 0.1, np.float64), k, Z, np.float64)
    # the GPU form is at least as close to FP64 as the FP32 restatement (up to 2x)
    assert rel(per.cpu().numpy(), o64) <= 2 * rel(o_vals, o64) + 1e-6


def test_nonpositive_ritz_value_gives_nan(cuda):
    A = np.diag(np.array([2.0, -1.0, 3.0, 0.5], np.float32))
    z = np.ones(4, np.float32)
    integrand = funm.monte_carlo_funm_sym_logdet(decomp.tridiag_sym(4))
    assert np.isnan(float(integrand(operators.dense_matvec, z, A)))


def test_deterministic_repeat(cuda):
    n, P, k = 1024, 32, 24
    A = torch.as_tensor(problems.dense_spd(n, seed=2), device=cuda)
    Z = stochtrace.sampler_signs(np.zeros(n), num=P).flat(prng.PRNGKey(2), cuda)
    a, _ = funm.logdet_per_probe(operators.dense_matvec, Z, k, (A,))
    b, _ = funm.logdet_per_probe(operators.dense_matvec, Z, k, (A,))
    assert torch.equal(a, b)


def test_profile_hooks_bracket_selected_classes(cuda):
    # measurement hooks: with only the matvec class selected, its k launches of
    # a depth-k solve are timed and no other class records events
    from matfree_amd import _runtime

    n, P, k = 300, 4, 6
    A = torch.as_tensor(sym_from_eigs(np.linspace(1.0, 3.0, n), 5), device=cuda)
    Z = oprng.rademacher(oprng.prng_key(2), (P, n), np.float32)
    ctx = _runtime.context(cuda)
    classes = ("matvec", "dots", "project_dots", "project", "normalize", "quadrature")
    ctx.profile(True, ("matvec",))
    for kc in classes:
        ctx.profile_read(kc)
    funm.logdet_per_probe(operators.dense_matvec, torch.as_tensor(Z, device=cuda), k, (A,))
    ms, cnt = ctx.profile_read("matvec")
    assert cnt == k and ms > 0.0
    for kc in classes[1:]:
        assert ctx.profile_read(kc) == (0.0, 0)
    ctx.profile(True)
    funm.logdet_per_probe(operators.dense_matvec, torch.as_tensor(Z, device=cuda), k, (A,))
    assert ctx.profile_read("matvec")[1] == k
    assert ctx.profile_read("quadrature")[1] == 1
    ctx.profile(False)
    funm.logdet_per_probe(operators.dense_matvec, torch.as_tensor(Z, device=cuda), k, (A,))
    assert ctx.profile_read("matvec") == (0.0, 0)


# ---------------------------------------------------------------------------
# full sizes: size-independent properties at the headline configurations
# ---------------------------------------------------------------------------


def test_full_size_rbf_logdet_within_1e3_of_fp64(cuda):
    # n = 32768 RBF (l = 0.5, s2 = 1e-2), 128 probes, depth 64, bf16x2: the
    # estimate is within 1e-3 of the FP64 dense-Cholesky log-determinant (the
    # bar the BF16 form is held to)
    from matfree_amd import problems as pr

    n, P, k = 32768, 128, 64
    X = torch.as_tensor(pr.rbf_points(n, seed=0), device=cuda)
    Z = stochtrace.sampler_signs(np.zeros(n, np.float32), num=P).flat(prng.PRNGKey(0), cuda)
    per, mean = funm.logdet_per_probe(operators.RBF(X, precision="bf16x2"), Z, k, (0.5, 1e-2))
    K = pr.rbf_dense64_device(X.cpu().numpy(), 0.5, 1e-2, cuda)
    exact = pr.logdet_cholesky_device(K)
    del K
    torch.cuda.empty_cache()
    assert abs(float(mean) - exact) <= 1e-3 * abs(exact)
    assert bool(torch.isfinite(per).all())


def test_full_size_dense_decomposition_properties(cuda):
    # n = 16384 stored dense A = B B^T / n + 0.1 I, FP32, depth 64 on 2 of the
    # 256 probes: Q Q^T = I and A Q^T = Q^T T + r e_k^T (tests/test_decomp/
    # test_tridiag_sym.py:62-65) to FP32 accuracy at the full dimension
    from matfree_amd import problems as pr

    n, k = 16384, 64
    A = pr.dense_spd_device(n, seed=0, shift=0.1, device=cuda)
    z = stochtrace.sampler_signs(np.zeros(n, np.float32), num=2).flat(prng.PRNGKey(0), cuda)[1]
    Q, T, r, _ = decomp.tridiag_sym(k)(operators.dense_matvec, z, A)
    Q64, T64, r64 = Q.double(), T.double(), r.double()
    eye = torch.eye(k, dtype=torch.float64, device=cuda)
    assert float(torch.linalg.matrix_norm(Q64 @ Q64.T - eye)) <= 1e-4
    lhs = A.double() @ Q64.T
    rhs = Q64.T @ T64
    rhs[:, -1] += r64
    assert float(torch.linalg.matrix_norm(lhs - rhs) / torch.linalg.matrix_norm(lhs)) <= 1e-4


@pytest.mark.parametrize("n", [2050, 2051, 2052])
def test_logdet_same_across_stream_widths(cuda, n):
    # the Gram-Schmidt kernels stream rows in aligned float4 groups (n % 4 == 0),
    # float4 groups shifted to each row's 16-byte boundary, float2 or single
    # floats depending on n; each reproduces the oracle on the same problem
    P, k = 6, 20
    A = problems.dense_spd(n, seed=3, shift=0.5)
    Z = oprng.rademacher(oprng.prng_key(11), (P, n), np.float32)
    per, _ = funm.logdet_per_probe(operators.dense_matvec, torch.as_tensor(Z, device=cuda), k,
                                   (torch.as_tensor(A, device=cuda),))
    o_vals, _, _ = oslq.quadform_logdet(oslq.DenseOperator(A), k, Z, np.float32)
    assert np.allclose(per.cpu().numpy(), o_vals, rtol=1e-5)
