"""PRNG keys in the layout of ``jax.random.PRNGKey`` (backend/prng.py:6-7).

A key is a pair of uint32 words ``[seed >> 32, seed & 0xffffffff]``; the
probe kernel consumes the two words directly.
"""

import numpy as np


def prng_key(seed):
    """``jax.random.PRNGKey(seed)`` for the default threefry2x32 implementation."""
    seed = int(seed)
    return np.array([(seed >> 32) & 0xFFFFFFFF, seed & 0xFFFFFFFF], dtype=np.uint32)


PRNGKey = prng_key


def key_words(key):
    """The two uint32 words of a key (array-like of length 2)."""
    arr = np.asarray(key if not hasattr(key, "cpu") else key.cpu().numpy()).astype(np.uint64)
    if arr.shape != (2,):
        raise ValueError(f"a PRNG key is a pair of uint32 words, got shape {arr.shape}")
    return int(arr[0]) & 0xFFFFFFFF, int(arr[1]) & 0xFFFFFFFF
