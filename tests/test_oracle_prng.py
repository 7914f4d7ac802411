"""Pins the oracle's probe generator (a1) against known answers.

* Random123 threefry2x32-20 known-answer vectors (the published KAT file of
  the algorithm JAX implements);
* matfree's README known answer (README.md:52-76): 10,000 Rademacher probes
  under PRNGKey(1) give the Hutchinson trace estimate 504.0 of A^T A with
  A = arange(12).reshape(6, 2); that value holds iff exactly 4980 probe rows
  have equal signs;
* the sampler properties of tests/test_stochtrace/test_samplers.py:46-77.
"""

import json
import os

import numpy as np
import pytest

from oracle import prng

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "prng_kat.json")


def _kat():
    with open(GOLDEN) as f:
        return json.load(f)


def test_threefry_known_answers():
    for case in _kat()["threefry2x32_20"]:
        key = [int(x, 16) for x in case["key"]]
        ctr = [int(x, 16) for x in case["ctr"]]
        y0, y1 = prng.threefry2x32(key, [ctr[0]], [ctr[1]])
        assert [int(y0[0]), int(y1[0])] == [int(x, 16) for x in case["out"]]


def test_prng_key_layout():
    assert prng.prng_key(1).tolist() == [0, 1]
    assert prng.prng_key(2**32 + 5).tolist() == [1, 5]


def test_readme_known_answer_504():
    kat = _kat()["readme_trace"]
    A = np.arange(12.0, dtype=np.float32).reshape(6, 2)
    M = A.T @ A
    z = prng.rademacher(prng.prng_key(kat["seed"]), (kat["num"], 2), np.float32)
    values = np.einsum("pi,ij,pj->p", z, M, z).astype(np.float32)
    assert int(np.sum(z[:, 0] == z[:, 1])) == kat["equal_sign_rows"]
    assert float(np.mean(values, dtype=np.float32)) == kat["estimate"]


def test_legacy_layout_is_rejected_by_the_readme_known_answer():
    # The non-partitionable layout gives a different stream; the README pins
    # the partitionable one.
    kat = _kat()["readme_trace"]
    bits = prng.random_bits32_legacy(prng.prng_key(kat["seed"]), (kat["num"], 2))
    z = np.where((bits >> np.uint32(31)) == 0, 1.0, -1.0)
    assert int(np.sum(z[:, 0] == z[:, 1])) != kat["equal_sign_rows"]


def test_rademacher_is_bit31_of_the_uniform_construction():
    # rademacher = 2 * (uniform < 0.5) - 1 with uniform from the top 23 bits
    key = prng.prng_key(7)
    bits = prng.random_bits32(key, (257,))
    uni = ((bits >> np.uint32(9)) | np.uint32(0x3F800000)).view(np.float32) - np.float32(1.0)
    expected = np.where(uni < 0.5, 1.0, -1.0).astype(np.float32)
    np.testing.assert_array_equal(prng.rademacher(key, (257,)), expected)


@pytest.mark.parametrize("n", [5])
def test_sampler_signs_properties(n):
    # tests/test_stochtrace/test_samplers.py:46-77
    num = 100_000
    sample = prng.sampler_signs(n, num)
    x = sample(prng.prng_key(1))
    assert np.allclose(np.abs(x), 1)
    assert np.allclose(np.mean(x), 0, atol=1e-2)
    assert np.allclose(np.cov(x, rowvar=False, ddof=0), np.eye(n), atol=1e-2)
    assert np.array_equal(x, sample(prng.prng_key(1)))
    assert not np.allclose(x, sample(prng.prng_key(2)))
