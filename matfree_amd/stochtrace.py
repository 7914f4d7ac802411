"""Stochastic trace estimation on the GPU (matfree.stochtrace subset).

* :func:`sampler_signs` (stochtrace.py:932-937): Rademacher probes generated
  by libslq's threefry kernel, bit-identical to ``jax.random.rademacher``;
* :func:`estimator_monte_carlo` (stochtrace.py:7-52): probe mean of an
  integrand, with every probe processed in one batched call;
* :func:`estimator_monte_carlo_mean_and_sem` (stochtrace.py:55-89): the mean
  and its standard error.
"""

import ctypes

import numpy as np
import torch

from . import _runtime, _tree, prng


class _SignSampler:
    """``sample(key)`` of ``sampler_signs(*args_like, num)``."""

    def __init__(self, args_like, num):
        if len(args_like) != 1:
            raise TypeError("sampler_signs expects exactly one args_like PyTree")
        self.args_like = args_like[0]
        leaves = _tree._leaves(self.args_like)
        for leaf in leaves:
            dt = leaf.dtype if hasattr(leaf, "dtype") else np.asarray(leaf).dtype
            if (isinstance(dt, torch.dtype) and dt.is_complex) or (
                not isinstance(dt, torch.dtype) and np.dtype(dt).kind == "c"
            ):
                raise NotImplementedError(
                    "complex probes (Steinhaus, stochtrace.py:967-977) are not part of matfree_amd"
                )
        self.n = int(sum(int(np.prod(getattr(leaf, "shape", np.shape(leaf)), dtype=np.int64))
                         for leaf in leaves))
        self.num = int(num)

    def flat(self, key, device=None):
        """The probes as one (num, n) float32 tensor."""
        device = device or _runtime.target_device()
        k0, k1 = prng.key_words(key)
        out = torch.empty((self.num, self.n), dtype=torch.float32, device=device)
        ctx = _runtime.context(device)
        ctx.call("slq_sampler_signs", k0, k1, self.num, self.n, _runtime.ptr(out))
        return out

    def __call__(self, key):
        flat = self.flat(key)
        if _tree.is_leaf(self.args_like):
            shape = tuple(getattr(self.args_like, "shape", np.shape(self.args_like)))
            return flat.reshape(self.num, *shape)
        _, unravel = _tree.ravel(self.args_like, flat.device)
        return unravel(flat)


def sampler_signs(*args_like, num):
    """Rademacher probes shaped like ``args_like`` (stochtrace.py:932-937, 957-964)."""
    return _SignSampler(args_like, num)


def estimator_monte_carlo_mean_and_sem(integrand, /, sampler):
    """``(mean, sem)`` of the integrand over the probes (stochtrace.py:55-89).

    ``sem = std(samples) / sqrt(num_samples)`` (population standard deviation),
    both reduced on the device in FP64.
    """
    if not hasattr(integrand, "batched"):
        raise TypeError(
            "matfree_amd's estimator runs integrands of matfree_amd.funm "
            f"(e.g. monte_carlo_funm_sym_logdet), not {type(integrand).__name__}"
        )

    def estimate(matvecs, key, *parameters):
        if isinstance(sampler, _SignSampler):
            samples = sampler.flat(key, _runtime.target_device(*parameters))
        else:
            samples = sampler(key)
        values = integrand.batched(matvecs, samples, *parameters)
        if (not isinstance(values, torch.Tensor) or values.ndim != 1
                or values.dtype != torch.float64 or not values.is_cuda):
            raise TypeError(
                "matfree_amd's mean-and-SEM estimator reduces scalar integrands (one float64 "
                "value per probe on the GPU, e.g. monte_carlo_funm_sym_logdet); "
                f"{type(integrand).__name__} returns "
                f"{tuple(values.shape) if isinstance(values, torch.Tensor) else type(values)}")
        values = values.contiguous()
        mean = torch.empty((), dtype=torch.float64, device=values.device)
        sem = torch.empty((), dtype=torch.float64, device=values.device)
        ctx = _runtime.context(values.device)
        ctx.call("slq_mean_and_sem", _runtime.ptr(values), values.shape[0], _runtime.ptr(mean),
                 _runtime.ptr(sem))
        return mean, sem

    return estimate


def estimator_monte_carlo(integrand, /, sampler):
    """Monte-Carlo estimator ``mean_p integrand(matvec, z_p, *params)`` (stochtrace.py:7-52)."""
    if not hasattr(integrand, "estimate"):
        raise TypeError(
            "matfree_amd's estimator runs integrands of matfree_amd.funm "
            f"(e.g. monte_carlo_funm_sym_logdet), not {type(integrand).__name__}"
        )

    def estimate(matvecs, key, *parameters):
        if isinstance(sampler, _SignSampler):
            samples = sampler.flat(key, _runtime.target_device(*parameters))
        else:
            samples = sampler(key)
        return integrand.estimate(matvecs, samples, *parameters)

    return estimate
