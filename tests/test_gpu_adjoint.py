"""GPU reverse mode of the SLQ log-determinant (run on an MI355X).

The HIP adjoint runs in FP32 (the FP64 tridiagonal eigen-cotangent aside), so
its cotangents are compared with the oracle's FP64 restatement of
decomp.py:480-600 + the eigh VJP (itself pinned by central differences and a
closed form in tests/test_oracle_slq.py), with the tolerance stated per test,
and with central differences of the GPU forward pass.
"""

import numpy as np
import pytest
import torch

from matfree_amd import decomp, funm, operators, problems, prng, stochtrace
from oracle import prng as oprng
from oracle import slq as oslq

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("n,P,k,dim", [(256, 6, 12, 3), (301, 5, 17, 3), (260, 4, 10, 4)])
def test_rbf_cotangents_match_oracle(cuda, n, P, k, dim):
    X = problems.rbf_points(n, seed=4, dim=dim)
    theta = (0.5, 0.1)
    Z = oprng.rademacher(oprng.prng_key(7), (P, n), np.float32)
    per, mean, grads = funm.logdet_value_and_grad(
        operators.RBF(X), torch.as_tensor(Z, device=cuda), k, theta
    )
    est64, g64 = oslq.slq_logdet_vjp(oslq.RBFOperator(X, *theta, np.float64), k, Z)
    g = np.array([float(grads[0]), float(grads[1])])
    # FP32 forward + FP32 adjoint against the FP64 restatement: 2e-4 normwise
    assert rel(g, g64) < 2e-4, (g, g64)
    assert abs(float(mean) - est64) <= 2e-5 * abs(est64)
    vals, _ = funm.logdet_per_probe(operators.RBF(X), torch.as_tensor(Z, device=cuda), k, theta)
    assert torch.equal(vals, per)


def test_rbf_cotangents_match_oracle_at_depth_32(cuda):
    # 32 probes, depth 32, (l, s2) = (0.5, 1e-2) as in the gradient
    # configuration, on n = 2048 points where the FP64 restatement is quick
    n, P, k = 2048, 32, 32
    X = problems.rbf_points(n, seed=0)
    theta = (0.5, 1e-2)
    Z = oprng.rademacher(oprng.prng_key(0), (P, n), np.float32)
    _, mean, grads = funm.logdet_value_and_grad(
        operators.RBF(X), torch.as_tensor(Z, device=cuda), k, theta
    )
    est64, g64 = oslq.slq_logdet_vjp(oslq.RBFOperator(X, *theta, np.float64), k, Z)
    g = np.array([float(grads[0]), float(grads[1])])
    assert rel(g, g64) < 1e-3, (g, g64)
    assert abs(float(mean) - est64) <= 1e-4 * abs(est64)


@pytest.mark.parametrize("precision,tol", [("bf16x3", 1e-3), ("bf16x2", 5e-3)])
def test_rbf_cotangents_bf16_forms(cuda, precision, tol):
    # the BF16 block matvec in the forward and reverse passes and the
    # split-BF16 parameter contraction, at the configuration of the depth-32
    # test, against the FP64 restatement (tolerance: the FP32 form's 1e-3, and
    # 5e-3 where the probes are rounded to BF16 once per matvec)
    n, P, k = 2048, 32, 32
    X = problems.rbf_points(n, seed=0)
    theta = (0.5, 1e-2)
    Z = oprng.rademacher(oprng.prng_key(0), (P, n), np.float32)
    _, mean, grads = funm.logdet_value_and_grad(
        operators.RBF(X, precision=precision), torch.as_tensor(Z, device=cuda), k, theta
    )
    est64, g64 = oslq.slq_logdet_vjp(oslq.RBFOperator(X, *theta, np.float64), k, Z)
    g = np.array([float(grads[0]), float(grads[1])])
    print(f"{precision}: gradient rel {rel(g, g64):.2e}, value rel "
          f"{abs(float(mean) - est64) / abs(est64):.2e}", flush=True)
    assert rel(g, g64) < tol, (g, g64)


def test_rbf_cotangents_match_central_differences(cuda):
    # the gradient configuration: n = 8192 points in the unit cube, 32 probes,
    # depth 32, (l, s2) = (0.5, 1e-2), against fourth-order central
    # differences of the FP32 GPU forward pass.  The forward value carries
    # FP32 Lanczos rounding (~1e-6 of |logdet| ~ 4e4), so the steps are
    # h = 1e-2 for l and 1e-4 for s2 and the tolerances 2% and 0.5%.
    n, P, k = 8192, 32, 32
    X = problems.rbf_points(n, seed=0)
    op = operators.RBF(X)
    Z = stochtrace.sampler_signs(np.zeros(n, np.float32), num=P).flat(prng.PRNGKey(0), cuda)
    theta = np.array([0.5, 1e-2])
    _, mean, grads = funm.logdet_value_and_grad(op, Z, k, tuple(theta))
    g = np.array([float(grads[0]), float(grads[1])])

    def f(t):
        _, m = funm.logdet_per_probe(op, Z, k, (float(t[0]), float(t[1])))
        return float(m)

    fd = []
    for e, h in zip(np.eye(2), (1e-2, 1e-4)):
        d = h * e
        fd.append((-f(theta + 2 * d) + 8 * f(theta + d) - 8 * f(theta - d) + f(theta - 2 * d))
                  / (12 * h))
    assert abs(g[0] - fd[0]) <= 2e-2 * abs(fd[0]), (g, fd)
    assert abs(g[1] - fd[1]) <= 5e-3 * abs(fd[1]), (g, fd)


def test_full_depth_matches_dense_closed_form(cuda):
    # at k = n the integrand is x^T log(A(theta)) x exactly (as in
    # tests/test_oracle_slq.py); its derivative in closed form via the
    # divided differences of log at the eigenvalues of A
    rng = np.random.default_rng(12)
    n = 10
    X = rng.uniform(size=(n, 2)).astype(np.float32)
    theta = (0.7, 0.3)
    z = oprng.rademacher(oprng.prng_key(5), (1, n), np.float32)
    _, _, grads = funm.logdet_value_and_grad(operators.RBF(X), torch.as_tensor(z, device=cuda),
                                             n, theta)
    op = oslq.RBFOperator(X.astype(np.float64), *theta, np.float64)
    A = op.dense()
    w, U = np.linalg.eigh(A)
    L = np.empty((n, n))
    for i in range(n):
        for j in range(n):
            L[i, j] = 1 / w[i] if i == j else (np.log(w[i]) - np.log(w[j])) / (w[i] - w[j])
    y = U.T @ z[0].astype(np.float64)
    d2 = np.sum((X[:, None].astype(np.float64) - X[None]) ** 2, -1)
    dA_dl = np.exp(-d2 / (2 * theta[0] ** 2)) * d2 / theta[0] ** 3
    for dA, gi in zip([dA_dl, np.eye(n)], grads):
        expected = np.sum(np.outer(y, y) * L * (U.T @ dA @ U))
        assert np.isclose(float(gi), expected, rtol=2e-4), (float(gi), expected)


def test_coincident_eigenvalues_cotangent_takes_the_ql_path(cuda):
    # W21+ (shifted by 2) from e_1 at full depth: T = A exactly, whose largest
    # eigenvalues pair up to ~1e-14; the eigen cotangent must come from the QL
    # kernel (the twisted eigenvectors of a pair are not orthogonal).  The
    # derivative of [log A]_00 along a symmetric direction D in closed form:
    # sum_ij y_i y_j L_ij (U^T D U)_ij, y = U^T e_1, L the divided differences
    # of log (1 / w_i where two eigenvalues coincide)
    n = 21
    a = np.abs(np.arange(n) - 10).astype(np.float64) + 2.0
    A = np.diag(a) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    z = np.zeros((1, n), np.float32)
    z[0, 0] = 1.0
    _, _, (G,) = funm.logdet_value_and_grad(
        operators.dense_matvec, torch.as_tensor(z, device=cuda), n,
        (torch.as_tensor(A.astype(np.float32), device=cuda),))
    w, U = np.linalg.eigh(A)
    L = np.empty((n, n))
    for i in range(n):
        for j in range(n):
            close = abs(w[i] - w[j]) <= 1e-8 * w[i]
            L[i, j] = 1 / w[i] if close else (np.log(w[i]) - np.log(w[j])) / (w[i] - w[j])
    y = U.T @ z[0].astype(np.float64)
    D = np.random.default_rng(3).standard_normal((n, n))
    D = D + D.T
    expected = np.sum(np.outer(y, y) * L * (U.T @ D @ U))
    got = float(np.sum(G.cpu().numpy().astype(np.float64) * D))
    assert np.isclose(got, expected, rtol=1e-4), (got, expected)


def test_estimator_autograd(cuda):
    # the drop-in for jax.grad over the estimator: torch parameters that
    # require grad make estimate(...) differentiable
    n, P, k = 512, 8, 16
    X = problems.rbf_points(n, seed=2)
    op = operators.RBF(X)
    tridiag = decomp.tridiag_sym(k)
    integrand = funm.monte_carlo_funm_sym_logdet(tridiag)
    sampler = stochtrace.sampler_signs(np.zeros(n, np.float32), num=P)
    estimate = stochtrace.estimator_monte_carlo(integrand, sampler)
    ls = torch.tensor(0.4, dtype=torch.float64, device=cuda, requires_grad=True)
    s2 = torch.tensor(0.05, dtype=torch.float64, device=cuda, requires_grad=True)
    val = estimate(op, prng.PRNGKey(3), ls, s2)
    val.backward()
    Z = sampler.flat(prng.PRNGKey(3), cuda)
    _, mean, grads = funm.logdet_value_and_grad(op, Z, k, (0.4, 0.05))
    assert float(val.detach()) == float(mean)
    assert float(ls.grad) == float(grads[0])
    assert float(s2.grad) == float(grads[1])
    # without requires_grad the estimate is a plain tensor
    plain = estimate(op, prng.PRNGKey(3), 0.4, 0.05)
    assert not plain.requires_grad and float(plain) == float(mean)


def test_adjoint_k0_raises(cuda):
    X = problems.rbf_points(16, seed=1)
    Z = torch.ones((2, 16), device=cuda)
    with pytest.raises(ValueError, match="= 0"):
        funm.logdet_value_and_grad(operators.RBF(X), Z, 0, (0.5, 0.1))


def test_dense_cotangent_matches_oracle(cuda):
    # matvec(v, A) = A v with A the parameter (tests/test_decomp/
    # test_tridiag_sym_adjoint.py differentiates w.r.t. the matrix too)
    n, P, k = 200, 4, 12
    A = problems.dense_spd(n, seed=5, shift=1.0)
    Z = oprng.rademacher(oprng.prng_key(2), (P, n), np.float32)
    _, mean, (G,) = funm.logdet_value_and_grad(
        operators.dense_matvec, torch.as_tensor(Z, device=cuda), k, (torch.as_tensor(A, device=cuda),)
    )
    est64, G64 = oslq.slq_logdet_vjp(oslq.DenseOperator(A, np.float64), k, Z)
    assert rel(G.cpu().numpy(), G64) < 2e-4
    assert abs(float(mean) - est64) <= 2e-5 * abs(est64)


def test_dense_cotangent_autograd_direction(cuda):
    # <dA, D> against a central difference of the GPU forward along a
    # symmetric direction D (FP32 forward: h = 1e-3, tolerance 1e-2)
    n, P, k = 300, 8, 20
    A = torch.as_tensor(problems.dense_spd(n, seed=6, shift=1.0), device=cuda)
    D = torch.randn((n, n), device=cuda, generator=torch.Generator(device=cuda).manual_seed(0))
    D = (D + D.T) / n
    integrand = funm.monte_carlo_funm_sym_logdet(decomp.tridiag_sym(k))
    sampler = stochtrace.sampler_signs(np.zeros(n, np.float32), num=P)
    estimate = stochtrace.estimator_monte_carlo(integrand, sampler)
    Ap = A.clone().requires_grad_(True)
    estimate(operators.dense_matvec, prng.PRNGKey(4), Ap).backward()
    h = 1e-3
    fd = (float(estimate(operators.dense_matvec, prng.PRNGKey(4), A + h * D))
          - float(estimate(operators.dense_matvec, prng.PRNGKey(4), A - h * D))) / (2 * h)
    got = float(torch.sum(Ap.grad * D))
    assert abs(got - fd) <= 1e-2 * abs(fd), (got, fd)


def test_autograd_with_some_parameters_plain(cuda):
    # a requires-grad lengthscale beside a plain float noise: the noise gets
    # no gradient and the lengthscale the same as the direct call
    n, P, k = 400, 6, 14
    X = problems.rbf_points(n, seed=8)
    op = operators.RBF(X)
    integrand = funm.monte_carlo_funm_sym_logdet(decomp.tridiag_sym(k))
    sampler = stochtrace.sampler_signs(np.zeros(n, np.float32), num=P)
    estimate = stochtrace.estimator_monte_carlo(integrand, sampler)
    ls = torch.tensor(0.45, dtype=torch.float32, device=cuda, requires_grad=True)
    estimate(op, prng.PRNGKey(1), ls, 3e-2).backward()
    Z = sampler.flat(prng.PRNGKey(1), cuda)
    _, _, grads = funm.logdet_value_and_grad(op, Z, k, (0.45, 3e-2))
    assert ls.grad.dtype == torch.float32
    assert float(ls.grad) == pytest.approx(float(grads[0]), rel=1e-6)
    # and the other way round: a plain lengthscale, a requires-grad noise
    s2 = torch.tensor(3e-2, dtype=torch.float64, device=cuda, requires_grad=True)
    estimate(op, prng.PRNGKey(1), 0.45, s2).backward()
    assert float(s2.grad) == float(grads[1])


@pytest.mark.parametrize("reortho", ["full", "none"])
def test_kept_and_recomputed_reverse_pass_agree(cuda, reortho):
    # the reverse pass from the kept forward state and the one that recomputes
    # the forward give the same bits; with the basis kept, the forward holds
    # the state until backward
    n, P, k = 700, 8, 18
    X = problems.rbf_points(n, seed=9)
    op = operators.RBF(X)
    integrand = funm.monte_carlo_funm_sym_logdet(decomp.tridiag_sym(k, reortho=reortho))
    sampler = stochtrace.sampler_signs(np.zeros(n, np.float32), num=P)
    estimate = stochtrace.estimator_monte_carlo(integrand, sampler)
    out = {}
    old = funm.REVERSE_POLICY
    try:
        for policy in ("keep", "recompute"):
            funm.REVERSE_POLICY = policy
            ls = torch.tensor(0.5, dtype=torch.float64, device=cuda, requires_grad=True)
            s2 = torch.tensor(0.1, dtype=torch.float64, device=cuda, requires_grad=True)
            val = estimate(op, prng.PRNGKey(2), ls, s2)
            saved = val.grad_fn.saved
            assert (saved is not None) == (policy == "keep")
            if saved is not None:
                assert saved.nbytes() >= 4 * k * P * n
            val.backward()
            out[policy] = (float(val.detach()), float(ls.grad), float(s2.grad))
    finally:
        funm.REVERSE_POLICY = old
    assert out["keep"] == out["recompute"]
    Z = sampler.flat(prng.PRNGKey(2), cuda)
    _, mean, grads = funm.logdet_value_and_grad(op, Z, k, (0.5, 0.1), reortho)
    assert out["keep"] == (float(mean), float(grads[0]), float(grads[1]))


def test_saved_forward_dense_matches_direct(cuda):
    n, P, k = 300, 4, 16
    A = torch.as_tensor(problems.dense_spd(n, seed=3, shift=1.0), device=cuda)
    Z = torch.as_tensor(oprng.rademacher(oprng.prng_key(6), (P, n), np.float32), device=cuda)
    per, mean, saved = funm.logdet_save(operators.dense_matvec, Z, k, (A,))
    (G1,) = funm.logdet_vjp_saved(operators.dense_matvec, saved, n, (A,))
    (G2,) = funm.logdet_vjp_saved(operators.dense_matvec, saved, n, (A,))  # reusable
    saved.close()
    per2, mean2, (G,) = funm.logdet_value_and_grad(operators.dense_matvec, Z, k, (A,))
    assert torch.equal(per, per2) and float(mean) == float(mean2)
    assert torch.equal(G1, G) and torch.equal(G2, G)


def test_workspace_query_matches_pool_growth(cuda):
    # the scratch slq_workspace_bytes reports is what a fresh context grows to
    # (within the 256-byte alignment of each buffer)
    from matfree_amd import _lib

    n, P, k = 1500, 16, 20
    X = torch.as_tensor(problems.rbf_points(n, seed=2), device=cuda)
    for entry, reortho, grad in [("logdet", "full", False), ("logdet_lanczos", "none", False),
                                 ("logdet_vjp", "full", True),
                                 ("logdet_lanczos_vjp", "none", True)]:
        op = operators.RBF(X, precision="bf16x2")
        desc, keep = op.descriptor(n, (0.5, 0.1), cuda)
        stream = torch.cuda.Stream(cuda)
        with torch.cuda.stream(stream):
            ctx = _lib.context(cuda.index, stream.cuda_stream)
            need = ctx.workspace_bytes(entry, desc, P, k)
            Z = torch.ones((P, n), device=cuda)
            if grad:
                funm.logdet_value_and_grad(op, Z, k, (0.5, 0.1), reortho)
            else:
                funm.logdet_per_probe(op, Z, k, (0.5, 0.1), reortho=reortho)
            held = ctx.scratch_bytes()
            stream.synchronize()
            ctx.release()
            assert ctx.scratch_bytes() == 0
        assert held <= need <= held + 256 * 64, (entry, held, need)


def test_workspace_growth_matches_pool_growth(cuda):
    # slq_workspace_growth: how far a context that already holds scratch grows
    # in a call, given what it holds (the pool keeps each buffer at the
    # largest size any call needed), over a sequence of forward and gradient
    # calls of different shapes on one context
    from matfree_amd import _lib

    calls = [(1500, 16, 20, False), (1200, 48, 24, True), (1700, 8, 12, False),
             (900, 64, 30, True), (1500, 16, 20, False)]
    stream = torch.cuda.Stream(cuda)
    with torch.cuda.stream(stream):
        ctx = _lib.context(cuda.index, stream.cuda_stream)
        for n, P, k, grad in calls:
            X = torch.as_tensor(problems.rbf_points(n, seed=n), device=cuda)
            op = operators.RBF(X, precision="bf16x2")
            desc, keep = op.descriptor(n, (0.5, 0.1), cuda)
            before = ctx.scratch_bytes()
            grow = ctx.workspace_growth("logdet_vjp" if grad else "logdet", desc, P, k)
            Z = torch.ones((P, n), device=cuda)
            if grad:
                funm.logdet_value_and_grad(op, Z, k, (0.5, 0.1))
            else:
                funm.logdet_per_probe(op, Z, k, (0.5, 0.1))
            after = ctx.scratch_bytes()
            assert after - before <= grow <= after - before + 256 * 64, (n, P, k, grad)
        stream.synchronize()
        ctx.release()
