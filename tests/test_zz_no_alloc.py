"""slq_logdet_noalloc (run on an MI355X): the allocation-free variant of
slq_logdet, called through the C ABI, gives bit-for-bit the results of
slq_logdet on the same inputs, on two calls in a row with the second smaller
than the first (the static scratch keeps the first call's contents), and a
call that needs more than the static scratch is rejected before anything is
queued.
"""

import ctypes

import numpy as np
import pytest
import torch

from matfree_amd import _lib, _runtime, operators, problems
from oracle import prng as oprng

pytestmark = pytest.mark.gpu


def make(kind, n, P, seed, device):
    if kind == "rbf":
        X = torch.as_tensor(problems.rbf_points(n, seed=seed), device=device)
        op, params = operators.RBF(X, precision="bf16x2"), (0.5, 1e-2)
    else:
        A = torch.as_tensor(problems.dense_spd(n, seed=seed, shift=0.5), device=device)
        op, params = operators.dense_matvec, (A,)
    Z = torch.as_tensor(oprng.rademacher(oprng.prng_key(seed), (P, n), np.float32), device=device)
    return op, params, Z


def call(entry, op, params, Z, k):
    """One call of ``entry`` through the C ABI; outputs copied to the host."""
    device = Z.device
    P, n = Z.shape
    desc, keep = op.descriptor(n, params, device)
    per = torch.full((P,), float("nan"), dtype=torch.float64, device=device)
    mean = torch.full((), float("nan"), dtype=torch.float64, device=device)
    alpha = torch.full((P, k), float("nan"), dtype=torch.float32, device=device)
    beta = torch.full((P, k - 1), float("nan"), dtype=torch.float32, device=device)
    ctx = _runtime.context(device)
    ctx.call(entry, ctypes.byref(desc), _runtime.ptr(Z), P, k, _runtime.ptr(per),
             _runtime.ptr(mean), _runtime.ptr(alpha), _runtime.ptr(beta))
    torch.cuda.synchronize()
    del keep
    return per.cpu().numpy(), float(mean), alpha.cpu().numpy(), beta.cpu().numpy()


# (kind, n, P, k) of the first call, then of the smaller second call; odd n
# and n not a multiple of 4 run the shifted streaming kernels, RBF bf16x2 the
# probe planes and point records
SEQUENCES = [
    [("rbf", 3001, 40, 24), ("dense", 1502, 8, 12)],
    [("dense", 4096, 64, 40), ("rbf", 1000, 16, 10)],
]


@pytest.mark.parametrize("seq", SEQUENCES, ids=["rbf-then-dense", "dense-then-rbf"])
def test_noalloc_matches_logdet_on_two_calls(cuda, seq):
    for i, (kind, n, P, k) in enumerate(seq):
        print(f"call {i}: {kind} n={n} P={P} k={k}", flush=True)
        op, params, Z = make(kind, n, P, 90 + i, cuda)
        ref = call("slq_logdet", op, params, Z, k)
        got = call("slq_logdet_noalloc", op, params, Z, k)
        assert np.all(np.isfinite(ref[0]))
        np.testing.assert_array_equal(got[0], ref[0])
        assert got[1] == ref[1]
        np.testing.assert_array_equal(got[2], ref[2])
        np.testing.assert_array_equal(got[3], ref[3])


def test_noalloc_rejects_a_call_beyond_the_static_scratch(cuda):
    # n = 32768, 128 probes, depth 64 needs about 1.2 GB of scratch
    n, P, k = 32768, 128, 64
    op, params, Z = make("rbf", n, P, 7, cuda)
    desc, keep = op.descriptor(n, params, cuda)
    per = torch.zeros((P,), dtype=torch.float64, device=cuda)
    mean = torch.zeros((), dtype=torch.float64, device=cuda)
    ctx = _runtime.context(cuda)
    with pytest.raises(ValueError, match="static scratch"):
        ctx.call("slq_logdet_noalloc", ctypes.byref(desc), _runtime.ptr(Z), P, k,
                 _runtime.ptr(per), _runtime.ptr(mean), None, None)
    torch.cuda.synchronize()
    assert float(per.abs().sum()) == 0.0  # nothing was queued
    del keep
    assert ctx.workspace_bytes("logdet", desc, P, k) > _lib.NOALLOC_SCRATCH_BYTES
