"""Packed low-precision storage of a stored matrix (run on an MI355X).

slq_pack_matrix must produce exactly the bytes (and tile scales) of the
oracle's model of the formats (oracle/lowprec.py: bfloat16 and OCP FP8 E4M3,
round to nearest even, power-of-two tile scales); the block matvec on the
packed planes must then equal the FP64 product of the represented matrix
with the BF16-rounded vectors up to FP32 accumulation (normwise 1e-5: the
BF16 products are exact in FP32), including the split-K and unaligned
paths; and the log-determinant on packed storage is compared with the FP32
stored-matrix form and with the FP64 log-determinant of the represented
matrix (tolerances stated per form below).
"""

import numpy as np
import pytest
import torch

from matfree_amd import funm, operators, problems
from oracle import lowprec as lp
from oracle import prng as oprng

pytestmark = pytest.mark.gpu

STORAGES = ["bf16x2", "bf16", "fp8x2", "fp8"]


def sample_matrix(n, seed):
    """Random symmetric entries over several magnitudes (one tile scaled down
    by 1e-3 and one tile of zeros), so per-tile scales differ."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)).astype(np.float32)
    A = (A + A.T) / 2
    A[:32, :32] *= np.float32(1e-3)
    A[32:64, 64:96] = 0.0
    A[64:96, 32:64] = 0.0
    return A



@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("n", [70, 301])
def test_pack_bytes_match_model(cuda, storage, n):
    A = sample_matrix(n, n)
    M = operators.pack(torch.as_tensor(A, device=cuda), storage)
    packed, scales, _ = lp.pack_model(A, storage)
    got = M.packed.cpu().numpy()
    assert got.shape == packed.shape
    mism = np.flatnonzero(got != packed)
    assert mism.size == 0, (mism[:10], got[mism[:10]], packed[mism[:10]])
    if scales is None:
        assert M.scales is None
    else:
        np.testing.assert_array_equal(M.scales.cpu().numpy(), scales)


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("n,P", [(301, 16), (256, 40), (2048, 128), (1000, 130)])
def test_packed_matvec_equals_represented_product(cuda, storage, n, P):
    # n = 301: unaligned rows; P = 40 / 130: probe blocks padded; n = 2048 with
    # 128 probes: split-K over column slices
    A = sample_matrix(n, 3)
    V = np.random.default_rng(4).standard_normal((P, n)).astype(np.float32)
    M = operators.pack(torch.as_tensor(A, device=cuda), storage)
    W = operators.packed_matvec(torch.as_tensor(V, device=cuda), M).cpu().numpy()
    R = lp.represented(A, storage)
    ref = lp.bf16_round(V).astype(np.float64) @ R.T
    err = np.linalg.norm(W - ref) / np.linalg.norm(ref)
    print(f"{storage} n={n} P={P}: normwise {err:.2e} vs the represented product", flush=True)
    assert err <= 1e-5
    # against the exact FP32 matrix: the storage and the BF16 vectors
    exact = V.astype(np.float64) @ A.astype(np.float64).T
    err_a = np.linalg.norm(W - exact) / np.linalg.norm(exact)
    bound = {"bf16x2": 4e-3, "bf16": 6e-3, "fp8x2": 6e-3, "fp8": 6e-2}[storage]
    assert err_a <= bound, err_a


# Storage error of the log-determinant on A = B B^T / n + 0.1 I (the stored
# dense SPD family), measured with FP64 slogdet of the represented matrix on
# the host: n = 2048: bf16x2 9e-7, bf16 5e-4, fp8x2 6.5e-5, fp8 5.7e-2.
# The SLQ estimate on packed storage must track the FP32 stored-matrix
# estimate on the same probes by that error plus the effect of rounding the
# Krylov vectors to BF16 before each matvec, which shifts every probe's value
# by about 3e-4 here (measured per-probe max against the FP32 form: bf16x2
# 6.8e-4, bf16 2.0e-3, fp8x2 6.1e-4, fp8 7.6e-2; the FP8 form is recorded,
# not held to 1e-3).
@pytest.mark.parametrize("storage,tol", [("bf16x2", 1.5e-3), ("bf16", 4e-3), ("fp8x2", 1.5e-3),
                                         ("fp8", 0.3)])
def test_packed_logdet_tracks_fp32(cuda, storage, tol):
    n, P, k = 2048, 32, 30
    A = problems.dense_spd(n, seed=0, shift=0.1)
    At = torch.as_tensor(A, device=cuda)
    Z = torch.as_tensor(oprng.rademacher(oprng.prng_key(5), (P, n), np.float32), device=cuda)
    ref, ref_mean = funm.logdet_per_probe(operators.dense_matvec, Z, k, (At,))
    M = operators.pack(At, storage)
    got, got_mean = funm.logdet_per_probe(operators.packed_matvec, Z, k, (M,))
    ref, got = ref.cpu().numpy(), got.cpu().numpy()
    d = np.abs(got - ref) / np.abs(ref)
    dm = abs(float(got_mean) - float(ref_mean)) / abs(float(ref_mean))
    print(f"{storage}: per-probe max rel {d.max():.2e}, mean rel {dm:.2e}", flush=True)
    assert np.all(np.isfinite(got))
    assert d.max() <= tol


def test_packed_rejects_bad_parameters(cuda):
    A = torch.eye(64, device=cuda)
    with pytest.raises(ValueError):
        operators.pack(A, "fp16")
    M = operators.pack(A, "fp8")
    with pytest.raises(ValueError):
        operators.packed_matvec(torch.ones(65, device=cuda), M)
    with pytest.raises(TypeError):
        operators.packed_matvec(torch.ones(64, device=cuda), A)
