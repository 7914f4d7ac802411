"""The compact exponent form of the split-BF16 RBF matvec (bf16, bf16x2).

The kernel forms the base-2 exponent of a tile from one BF16 MFMA: the scaled
coordinates y = s x (s = sqrt(log2 e / (2 l^2))) rounded to two BF16 parts
y~ = y_h + y_l, the cross products 2 y~_i . y~_m exact, and each squared norm
|y~|^2 (FP32) as two BF16 parts.  Its tiles are therefore those of

    K~_im = exp2(2 y~_i . y~_m - N~_i - N~_m) + noise [i == m],
    N~ = the two-part BF16 rounding of the FP32 |y~|^2,

which is the RBF kernel of the points y~ (moved by at most 2^-18 of their
scaled coordinates) times D_i D_m with D = exp2(|y~|^2 - N~): a symmetric
positive definite matrix.  `compact_model` restates that arithmetic on the
host in FP64.  The CPU test checks the claim that its log-determinant is the
exact one to ~1e-9 relative; the GPU test checks the kernel against the model
with probes of +-1 (exact in BF16), where only the hi + lo representation of
the entries (2^-18 relative each) and the exponents' sums remain.  Both sides
of the kernel take the points relative to point 0, and so does the model.
"""

import numpy as np
import pytest
import torch

from matfree_amd import operators, problems

LOG2E = 1.4426950408889634


def _bf16(a):
    """Round FP32 values to BF16 (nearest even) and back to FP32."""
    t = torch.as_tensor(np.asarray(a, np.float32))
    return t.to(torch.bfloat16).to(torch.float32).numpy()


def compact_model(points, lengthscale, noise):
    """FP64 matrix of the compact exponent form (see the module docstring)."""
    X = np.zeros((len(points), 3), np.float32)
    X[:, : points.shape[1]] = points - points[0]  # FP32, relative to point 0
    s = np.float32(np.sqrt(LOG2E / (2.0 * lengthscale * lengthscale)))
    y = (s * X).astype(np.float32)
    yh = _bf16(y)
    yl = _bf16((y - yh).astype(np.float32))
    yt = yh.astype(np.float64) + yl.astype(np.float64)  # exact in FP32
    # FP32 |y~|^2 by three fused multiply-adds (each rounded once)
    nrm = np.zeros(len(X), np.float64)
    for d in range(3):
        nrm = np.float32(yt[:, d] * yt[:, d] + nrm).astype(np.float64)
    nh = _bf16((-nrm).astype(np.float32)).astype(np.float64)
    nl = _bf16((-nrm - nh).astype(np.float32)).astype(np.float64)
    N = -(nh + nl)
    arg = 2.0 * yt @ yt.T - N[:, None] - N[None, :]
    K = np.exp2(arg)
    K[np.diag_indices(len(X))] += noise
    return K


def _logdet(K):
    return 2.0 * np.sum(np.log(np.diag(np.linalg.cholesky(K))))


def test_compact_form_logdet_matches_exact_kernel():
    # n = 2048 points in the unit cube, l = 0.5, noise 1e-2: the perturbation
    # of the points and the norms' rounding move the log-determinant by ~1e-9
    # relative (measured), far inside the 1e-3 bar and the split entries'
    # 2^-17; an exponent with a dropped term or a wrong sign would not be
    X = problems.rbf_points(2048, seed=5)
    exact = _logdet(problems.rbf_dense64(X, 0.5, 1e-2))
    compact = _logdet(compact_model(X, 0.5, 1e-2))
    assert abs(compact - exact) <= 1e-7 * abs(exact)
    # the model is not the exact matrix (the check above is not vacuous)
    assert np.max(np.abs(compact_model(X, 0.5, 1e-2) - problems.rbf_dense64(X, 0.5, 1e-2))) > 1e-7


def _split_entries(K):
    """hi + lo BF16 representation of the FP32-rounded entries (bf16x2)."""
    Kf = K.astype(np.float32)
    hi = _bf16(Kf)
    return hi.astype(np.float64) + _bf16((Kf - hi).astype(np.float32))


def test_compact_model_bound_discriminates():
    # Points in [0, 8]^3 (relative to point 0: squared scaled coordinates up
    # to ~170): the GPU test's bound of 2^-13 (normwise, probes of +-1) sits
    # above what the hi + lo entries leave (~3e-6) and what the matrix core's
    # sums of the 16 slot products add at this range (up to 5.3e-5 measured;
    # summing them in sequence in FP32 gives several times less) and below the
    # distance of the model to the exact kernel
    # and to the model with a planted mistake (a dropped y_l y_l product, or
    # one-part norms), so a kernel that met the bound against the wrong
    # matrix would fail it.
    n = 700
    X = (problems.rbf_points(n, seed=6) * 8.0).astype(np.float32)
    K = compact_model(X, 0.5, 1e-2)
    V = np.random.default_rng(1).choice([-1.0, 1.0], size=(8, n))

    def dist(M):
        return np.linalg.norm(V @ M - V @ K) / np.linalg.norm(V @ K)

    s = np.float32(np.sqrt(LOG2E / (2.0 * 0.25)))
    y = (s * (X - X[0]).astype(np.float32)).astype(np.float32)
    yh = _bf16(y)
    yl = _bf16((y - yh).astype(np.float32)).astype(np.float64)
    eye = 1e-2 * np.eye(n)
    dropped_ll = (K - eye) * np.exp2(-2.0 * yl @ yl.T) + eye
    yt = yh.astype(np.float64) + yl
    N = np.float32(np.sum(yt * yt, axis=1)).astype(np.float64)
    N1 = -_bf16((-N).astype(np.float32)).astype(np.float64)  # one BF16 part
    one_part = np.exp2(2.0 * yt @ yt.T - N1[:, None] - N1[None, :]) + eye
    assert dist(_split_entries(K)) < 2.0**-13 / 8
    assert dist(problems.rbf_dense64(X, 0.5, 1e-2)) > 2.0**-13
    for wrong in (dropped_ll, one_part):
        assert dist(wrong) > 2.0**-13 * 1.5


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["bf16x2", "bf16"])
@pytest.mark.parametrize("n,P,dim", [(1000, 16, 3), (2053, 128, 3), (777, 40, 2),
                                     (300, 70, 1), (4099, 33, 3)])
def test_compact_exponent_matvec_matches_model(cuda, precision, n, P, dim):
    # points in [0, 8]^dim, where the model stands apart from the exact kernel
    X = (problems.rbf_points(n, seed=n + dim, dim=dim) * 8.0).astype(np.float32)
    V = np.random.default_rng(P).choice([-1.0, 1.0], size=(P, n)).astype(np.float32)
    W = operators.RBF(X, precision=precision)(torch.as_tensor(V, device=cuda), 0.5, 1e-2)
    W = W.cpu().numpy().astype(np.float64)
    want = V.astype(np.float64) @ compact_model(X, 0.5, 1e-2)
    exact = V.astype(np.float64) @ problems.rbf_dense64(X, 0.5, 1e-2)
    err = np.linalg.norm(W - want) / np.linalg.norm(want)
    err_exact = np.linalg.norm(W - exact) / np.linalg.norm(exact)
    print(f"{precision} n={n} P={P} dim={dim}: vs model {err:.2e}, vs exact {err_exact:.2e}",
          flush=True)
    # bf16x2: hi + lo entries (2^-18 relative each) and the exponents' sums on
    # the matrix core; bf16: entries rounded once (2^-9 relative each)
    if precision == "bf16x2":
        assert err <= 2.0**-13
        assert err < 0.5 * err_exact  # the kernel's matrix is the model's (4-7x closer)
    else:
        assert err <= 2.0**-8
