"""Serving loop (matfree_amd.serving): requests that overlap on the device
compute exactly what they compute alone, and the HBM admission rule holds.

The per-probe log-determinants of every request are compared bit for bit:
each request has its own stream and library context, and every kernel of the
path reduces in a fixed order, so running beside other requests must not
change a single bit.
"""

import numpy as np
import pytest
import torch

from matfree_amd import funm, operators, problems, prng, serving, stochtrace

pytestmark = pytest.mark.gpu


def mixed_requests(device):
    shapes = [("rbf", 3000, 32, 20), ("dense", 700, 8, 12), ("rbf", 1500, 4, 16),
              ("dense", 2048, 64, 24), ("rbf", 5000, 16, 30), ("dense", 300, 16, 9),
              ("rbf", 800, 128, 10), ("dense", 1200, 32, 18)]
    reqs = []
    for r, (kind, n, P, k) in enumerate(shapes):
        if kind == "rbf":
            X = torch.as_tensor(problems.rbf_points(n, seed=50 + r), device=device)
            mv, params = operators.RBF(X, precision="bf16x2"), (0.5, 1e-2)
        else:
            A = torch.as_tensor(problems.dense_spd(n, seed=50 + r, shift=0.5), device=device)
            mv, params = operators.dense_matvec, (A,)
        reqs.append(serving.Request(mv, params, n, P, k, seed=r, arrival=0.0))
    torch.cuda.synchronize()
    return reqs


def alone(req, device):
    Z = stochtrace.sampler_signs(np.zeros(req.n, np.float32), num=req.P).flat(
        prng.PRNGKey(req.seed), device)
    per, _ = funm.logdet_per_probe(req.matvec, Z, req.k, req.params)
    return per.cpu().numpy()


def test_results_alone_and_in_the_mix_are_identical(cuda):
    reqs = mixed_requests(cuda)
    ref = [alone(q, cuda) for q in reqs]
    out = serving.Server(cuda, slots=4).run(reqs)
    assert max(o.slot for o in out) > 0  # several slots were used
    overlap = any(a.start < b.finish and b.start < a.finish
                  for i, a in enumerate(out) for b in out[i + 1:])
    assert overlap  # requests were in flight together
    for o, r in zip(out, ref):
        np.testing.assert_array_equal(o.per_probe.cpu().numpy(), r)
        assert o.mean == pytest.approx(float(np.mean(r)), rel=1e-12)


def test_hbm_budget_serialises_requests(cuda):
    reqs = mixed_requests(cuda)[:5]
    # no two requests fit together (each needs at least the smallest footprint);
    # the larger ones are admitted because nothing else runs
    budget = min(serving.footprint(q) for q in reqs)
    out = serving.Server(cuda, slots=4, hbm_budget=budget).run(reqs)
    spans = sorted((o.start, o.finish) for o in out)
    for (s0, f0), (s1, f1) in zip(spans, spans[1:]):
        assert s1 >= f0  # the next one starts after the previous one was seen done
    for o, q in zip(out, reqs):
        np.testing.assert_array_equal(o.per_probe.cpu().numpy(), alone(q, cuda))


def test_arrivals_are_respected(cuda):
    reqs = [q._replace(arrival=0.02 * i) for i, q in enumerate(mixed_requests(cuda)[:4])]
    out = serving.Server(cuda, slots=2).run(reqs)
    for o, q in zip(out, reqs):
        assert o.start >= q.arrival and o.finish >= o.start
