"""GPU funm_lanczos_sym and the mean/SEM estimator (run on an MI355X).

funm_lanczos_sym (funm.py:114-147) shares the forward kernels with the
log-determinant path; the new parts are the FP64 f(T) e_1 and the final
combination of basis vectors.  Tolerances are stated per test.
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


def test_reference_configuration_sin(cuda):
    # tests/test_funm/test_funm_lanczos_sym.py:8-36: n = 11, eigenvalues in
    # [0.01, 0.99], f = sin, tridiag_sym(6), atol 1e-6
    n = 11
    A = sym_from_eigs(np.linspace(0.01, 0.99, n), 1)
    v = np.random.default_rng(2).standard_normal(n).astype(np.float32)
    matfun_vec = funm.funm_lanczos_sym(funm.dense_funm_sym_eigh(np.sin), decomp.tridiag_sym(6))
    got = matfun_vec(operators.dense_matvec, [(v,)], A)
    [(received,)] = got
    w, U = np.linalg.eigh(A.astype(np.float64))
    expected = (U * np.sin(w)) @ U.T @ v.astype(np.float64)
    assert np.allclose(received.cpu().numpy(), expected, atol=1e-6)


@pytest.mark.parametrize("fun,name", [(np.exp, "exp"), (np.log, "log"), (np.sqrt, "sqrt"),
                                      (np.cos, "cos"), (np.reciprocal, "reciprocal")])
def test_batched_matches_oracle(cuda, fun, name):
    n, P, k = 300, 7, 15
    A = problems.dense_spd(n, seed=8, shift=1.0)
    V = np.random.default_rng(3).standard_normal((P, n)).astype(np.float32)
    f = funm.funm_lanczos_sym(funm.dense_funm_sym_eigh(fun), decomp.tridiag_sym(k))
    got = f.batched(operators.dense_matvec, torch.as_tensor(V, device=cuda),
                    torch.as_tensor(A, device=cuda)).cpu().numpy()
    o32 = oslq.funm_lanczos_sym(oslq.DenseOperator(A), k, V, fun, np.float32)
    o64 = oslq.funm_lanczos_sym(oslq.DenseOperator(A, np.float64), k, V, fun, np.float64)
    # the FP32 GPU path is as close to FP64 as the FP32 restatement (up to 2x)
    assert rel(got, o64) <= 2 * rel(o32, o64) + 1e-6, (name, rel(got, o64), rel(o32, o64))


def test_full_depth_exact_rbf(cuda):
    # k = n: |v| Q^T f(T) e1 = f(A) v exactly (FP32 path: 1e-4 normwise)
    n = 40
    X = problems.rbf_points(n, seed=5)
    v = np.random.default_rng(4).standard_normal((2, n)).astype(np.float32)
    f = funm.funm_lanczos_sym(funm.dense_funm_sym_eigh(np.exp), decomp.tridiag_sym(n))
    got = f.batched(operators.RBF(X), torch.as_tensor(v, device=cuda), 0.4, 0.3).cpu().numpy()
    A = oslq.RBFOperator(X, 0.4, 0.3, np.float64).dense()
    w, U = np.linalg.eigh(A)
    expected = v.astype(np.float64) @ ((U * np.exp(w)) @ U.T)
    assert rel(got, expected) < 1e-4


def test_log_quadform_matches_integrand(cuda):
    # z^T (|z| Q^T log(T) e1) = |z|^2 [log T]_00, the logdet integrand
    n, P, k = 512, 8, 20
    A = torch.as_tensor(problems.dense_spd(n, seed=9, shift=1.0), device=cuda)
    Z = stochtrace.sampler_signs(np.zeros(n), num=P).flat(prng.PRNGKey(6), cuda)
    f = funm.funm_lanczos_sym(funm.dense_funm_sym_eigh(np.log), decomp.tridiag_sym(k))
    fz = f.batched(operators.dense_matvec, Z, A)
    per, _ = funm.logdet_per_probe(operators.dense_matvec, Z, k, (A,))
    quad = (Z.double() * fz.double()).sum(dim=1)
    assert torch.allclose(quad, per, rtol=1e-5)


def w21_plus():
    # Wilkinson's W21+ shifted by 2: its largest eigenvalues come in pairs that
    # agree to ~1e-14, below what bisection separates
    n = 21
    a = np.abs(np.arange(n) - 10).astype(np.float64) + 2.0
    return np.diag(a) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)


@pytest.mark.parametrize("fun", [np.log, np.exp])
def test_coincident_eigenvalues_take_the_ql_path(cuda, fun):
    # from e_1 and e_21 the Krylov tridiagonal is W21+ itself (exact in FP32);
    # the twisted eigenvectors of its coincident pairs are not orthogonal, so
    # the four-wave kernel must hand the probes to the QL kernel: f(A) e_1 and
    # f(A) e_21 in closed form (FP32 output: 1e-6 relative)
    A = w21_plus()
    n = A.shape[0]
    V = np.zeros((2, n), np.float32)
    V[0, 0] = 1.0
    V[1, -1] = 1.0
    f = funm.funm_lanczos_sym(funm.dense_funm_sym_eigh(fun), decomp.tridiag_sym(n))
    got = f.batched(operators.dense_matvec, torch.as_tensor(V, device=cuda),
                    torch.as_tensor(A.astype(np.float32), device=cuda)).cpu().numpy()
    w, U = np.linalg.eigh(A)
    expected = V.astype(np.float64) @ ((U * fun(w)) @ U.T)
    assert rel(got, expected) < 1e-6, rel(got, expected)


def test_unsupported_matfun_raises(cuda):
    with pytest.raises(NotImplementedError):
        funm.funm_lanczos_sym(funm.dense_funm_sym_eigh(np.tanh), decomp.tridiag_sym(4))


def test_mean_and_sem(cuda):
    n, P, k = 1024, 40, 24
    A = torch.as_tensor(problems.dense_spd(n, seed=2, shift=1.0), device=cuda)
    integrand = funm.monte_carlo_funm_sym_logdet(decomp.tridiag_sym(k))
    sampler = stochtrace.sampler_signs(np.zeros(n, np.float32), num=P)
    mean, sem = stochtrace.estimator_monte_carlo_mean_and_sem(integrand, sampler)(
        operators.dense_matvec, prng.PRNGKey(3), A)
    per, m2 = funm.logdet_per_probe(operators.dense_matvec, sampler.flat(prng.PRNGKey(3), cuda),
                                    k, (A,))
    o_mean, o_sem = oslq.estimator_mean_and_sem(per.cpu().numpy(), np.float64)
    assert abs(float(mean) - o_mean) <= 1e-12 * abs(o_mean)
    assert abs(float(sem) - o_sem) <= 1e-10 * abs(o_sem)
    assert float(mean) == pytest.approx(float(m2), rel=1e-14)


def test_mean_and_sem_rejects_vector_integrands(cuda):
    # the device reduction takes one float64 value per probe
    n = 64
    A = torch.as_tensor(problems.dense_spd(n, seed=1, shift=1.0), device=cuda)
    f = funm.funm_lanczos_sym(funm.dense_funm_sym_eigh(np.exp), decomp.tridiag_sym(5))
    sampler = stochtrace.sampler_signs(np.zeros(n, np.float32), num=3)
    with pytest.raises(TypeError, match="scalar integrands"):
        stochtrace.estimator_monte_carlo_mean_and_sem(f, sampler)(operators.dense_matvec,
                                                                  prng.PRNGKey(0), A)
