"""Restatement of the probe generator on matfree's hot path -- TEST INFRASTRUCTURE.

matfree draws probes with ``stochtrace.sampler_signs`` (stochtrace.py:932-937)
-> ``_sampler_from_jax_random`` (stochtrace.py:957-964) -> ``_uniform_signs``
(stochtrace.py:974-977) -> ``prng.rademacher`` (backend/prng.py:26-29), i.e.
``jax.random.rademacher(key, (num, n), float32)``.  The arithmetic lives in JAX
(third-party, unpinned ``jax[cpu]`` in pyproject.toml:29, not vendored).  The
published algorithm restated here:

* key: ``PRNGKey(seed)`` -> uint32 ``[seed >> 32, seed & 0xffffffff]``
  (backend/prng.py:6-7);
* threefry2x32 with 20 rounds (Salmon et al., SC'11 / Random123);
* 32-bit random bits in the "partitionable" layout: element with flat row-major
  index ``m`` is encrypted from counter ``(m >> 32, m & 0xffffffff)`` and its
  bits are ``y0 ^ y1``;
* ``rademacher = 2 * (uniform < 0.5) - 1`` with
  ``uniform = bitcast((bits >> 9) | 0x3f800000) - 1``, i.e. ``+1`` exactly
  when bit 31 of the word is clear.

Pinned by the Random123 known-answer vectors and by matfree's README known
answer (README.md:52-76): with ``PRNGKey(1)`` and 10,000 probes of length 2,
the trace estimate of ``A^T A`` (``A = arange(12).reshape(6, 2)``) is exactly
``504.0``, which holds iff exactly 4980 rows have equal signs.  The legacy
(non-partitionable) layout gives 5029 rows (508.9) and is rejected by that test.
"""

import numpy as np

_ROT = ((13, 15, 26, 6), (17, 29, 16, 24))
_PARITY = np.uint32(0x1BD11BDA)


def prng_key(seed):
    """``jax.random.PRNGKey(seed)`` as used at backend/prng.py:6-7."""
    seed = int(seed)
    return np.array([(seed >> 32) & 0xFFFFFFFF, seed & 0xFFFFFFFF], dtype=np.uint32)


def _rotl(x, r):
    return (x << np.uint32(r)) | (x >> np.uint32(32 - r))


def threefry2x32(key, x0, x1):
    """threefry2x32-20 on uint32 arrays ``x0, x1`` under ``key = (k0, k1)``."""
    k0 = np.uint32(key[0])
    k1 = np.uint32(key[1])
    ks = (k0, k1, np.uint32(k0 ^ k1 ^ _PARITY))
    x0 = np.array(x0, dtype=np.uint32, copy=True)
    x1 = np.array(x1, dtype=np.uint32, copy=True)
    with np.errstate(over="ignore"):
        x0 += ks[0]
        x1 += ks[1]
        for group in range(5):
            for r in _ROT[group % 2]:
                x0 += x1
                x1 = _rotl(x1, r)
                x1 ^= x0
            x0 += ks[(group + 1) % 3]
            x1 += ks[(group + 2) % 3] + np.uint32(group + 1)
    return x0, x1


def random_bits32(key, shape):
    """32-bit words in JAX's partitionable threefry layout."""
    size = int(np.prod(shape, dtype=np.int64))
    m = np.arange(size, dtype=np.uint64)
    hi = (m >> np.uint64(32)).astype(np.uint32)
    lo = (m & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    y0, y1 = threefry2x32(key, hi, lo)
    return (y0 ^ y1).reshape(shape)


def random_bits32_legacy(key, shape):
    """32-bit words in JAX's original (non-partitionable) layout.

    Kept only so that a test can show the README known answer rejects it.
    """
    size = int(np.prod(shape, dtype=np.int64))
    count = np.arange(size, dtype=np.uint32)
    if size % 2:
        count = np.concatenate([count, np.zeros(1, np.uint32)])
    half = count.size // 2
    y0, y1 = threefry2x32(key, count[:half], count[half:])
    return np.concatenate([y0, y1])[:size].reshape(shape)


def rademacher(key, shape, dtype=np.float32):
    """``jax.random.rademacher(key, shape, dtype)`` (backend/prng.py:26-29)."""
    bits = random_bits32(key, shape)
    return np.where((bits >> np.uint32(31)) == 0, 1.0, -1.0).astype(dtype)


def sampler_signs(n, num, dtype=np.float32):
    """``stochtrace.sampler_signs(zeros(n), num=num)`` (stochtrace.py:932-964)."""

    def sample(key):
        return rademacher(key, (num, n), dtype)

    return sample
