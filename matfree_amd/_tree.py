"""Minimal ``ravel_pytree`` for the containers matfree's tests use.

matfree flattens vector PyTrees with ``jax.flatten_util.ravel_pytree``
(decomp.py:381-384, funm.py:227, stochtrace.py:958).  Supported here: arrays,
and (nested) lists, tuples and dicts of arrays; dict leaves are ordered by
sorted key, as in JAX.
"""

import numpy as np
import torch


def _leaves(tree):
    if isinstance(tree, dict):
        out = []
        for k in sorted(tree):
            out.extend(_leaves(tree[k]))
        return out
    if isinstance(tree, (list, tuple)):
        out = []
        for item in tree:
            out.extend(_leaves(item))
        return out
    return [tree]


def _rebuild(tree, it):
    if isinstance(tree, dict):
        return {k: _rebuild(tree[k], it) for k in sorted(tree)}
    if isinstance(tree, list):
        return [_rebuild(x, it) for x in tree]
    if isinstance(tree, tuple):
        return tuple(_rebuild(x, it) for x in tree)
    return next(it)


def is_leaf(tree):
    return not isinstance(tree, (dict, list, tuple))


def ravel(tree, device):
    """Flatten ``tree`` into one float32 tensor; return (flat, unravel)."""
    leaves = _leaves(tree)
    shapes, tensors = [], []
    for leaf in leaves:
        t = leaf if isinstance(leaf, torch.Tensor) else torch.as_tensor(np.asarray(leaf))
        shapes.append(tuple(t.shape))
        # not detached: autograd reaches the leaves through the flattening
        tensors.append(t.to(device=device, dtype=torch.float32).reshape(-1))
    flat = torch.cat(tensors) if len(tensors) > 1 else tensors[0].contiguous()
    sizes = [int(np.prod(s, dtype=np.int64)) for s in shapes]

    def unravel(vec):
        """Inverse of the flattening for a trailing axis of length n."""
        lead = tuple(vec.shape[:-1])
        parts, start = [], 0
        for size, shape in zip(sizes, shapes):
            parts.append(vec[..., start : start + size].reshape(*lead, *shape))
            start += size
        if is_leaf(tree):
            return parts[0]
        return _rebuild(tree, iter(parts))

    return flat, unravel
