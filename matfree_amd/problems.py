"""Seeded synthetic workloads of the SLQ log-determinant benchmarks.

* ``dense_spd(n, seed, shift)``: ``A = B B^T / n + shift I`` with ``B`` iid
  N(0, 1) in float32 (NumPy PCG64 stream of ``seed``), formed in FP64 and
  rounded once to float32;
* ``rbf_points(n, seed, dim)``: points uniform in ``[0, 1]^dim`` (float32).

Host generation is used for every size; the ``device`` variant of
``dense_spd`` forms the product on the GPU for large n.
"""

import numpy as np
import torch


def normal_matrix(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, n), dtype=np.float32)


def dense_spd(n, seed=0, shift=1.0):
    """Host float32 ``B B^T / n + shift I``."""
    B = normal_matrix(n, seed).astype(np.float64)
    A = B @ B.T / n
    A[np.diag_indices(n)] += shift
    return A.astype(np.float32)


def dense_spd_device(n, seed=0, shift=1.0, device="cuda"):
    """Same matrix as :func:`dense_spd` up to FP64 summation order, formed on the GPU."""
    B = torch.as_tensor(normal_matrix(n, seed), device=device).to(torch.float64)
    A = (B @ B.T) / n
    del B
    A.diagonal().add_(shift)
    return A.to(torch.float32)


def rbf_points(n, seed=0, dim=3):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, size=(n, dim)).astype(np.float32)


def rbf_dense64(points, lengthscale, noise):
    """Dense FP64 RBF matrix ``exp(-|x_i - x_j|^2 / (2 l^2)) + noise I`` (host)."""
    X = np.asarray(points, np.float64)
    sq = np.sum(X * X, axis=1)
    d2 = np.maximum(sq[:, None] + sq[None, :] - 2.0 * X @ X.T, 0.0)
    K = np.exp(-d2 / (2.0 * lengthscale**2))
    K[np.diag_indices(len(X))] += noise
    return K


def rbf_dense64_device(points, lengthscale, noise, device="cuda"):
    """Dense FP64 RBF matrix on the GPU (direct differences, no cancellation)."""
    X = torch.as_tensor(np.asarray(points), device=device).to(torch.float64)
    n = X.shape[0]
    K = torch.empty((n, n), dtype=torch.float64, device=device)
    rows = max(1, (1 << 27) // max(n, 1))
    for s in range(0, n, rows):
        e = min(n, s + rows)
        d2 = ((X[s:e, None, :] - X[None, :, :]) ** 2).sum(-1)
        K[s:e] = torch.exp(-d2 / (2.0 * lengthscale**2))
    K.diagonal().add_(noise)
    return K


def logdet_cholesky_device(A64):
    """FP64 Cholesky log-determinant of a symmetric positive-definite tensor."""
    L = torch.linalg.cholesky(A64)
    return float(2.0 * torch.log(torch.diagonal(L)).sum())
