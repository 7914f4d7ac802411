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
    # n = 3001 and 1502 (not multiples of 4) run the Gram-Schmidt passes that
    # start each row's float4 groups at its 16-byte boundary
    shapes = [("rbf", 3000, 32, 20), ("dense", 700, 8, 12), ("rbf", 1500, 4, 16),
              ("dense", 2048, 64, 24), ("rbf", 5000, 16, 30), ("dense", 300, 16, 9),
              ("rbf", 800, 128, 10), ("dense", 1200, 32, 18), ("rbf", 3001, 16, 22),
              ("dense", 1502, 32, 14)]
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


def c5_requests(device, rng):
    """A short stream from the serving distribution: n log-uniform in
    [2048, 32768], probes in {16, 32, 64, 128}, depth uniform in [16, 64], RBF
    in bf16x2 or a stored SPD matrix (n <= 16384), plus the two largest shapes:
    an n = 32768, 128-probe, depth-64 RBF request and a stored n = 16384 one;
    every third request asks for the gradient too."""
    shapes = [("rbf", 32768, 128, 64), ("dense", 16384, 64, 40)]
    for _ in range(8):
        n = int(np.exp(rng.uniform(np.log(2048), np.log(32768))))
        kind = "rbf" if (n > 16384 or rng.uniform() < 0.5) else "dense"
        shapes.append((kind, n, int(rng.choice([16, 32, 64, 128])), int(rng.integers(16, 65))))
    reqs = []
    for r, (kind, n, P, k) in enumerate(shapes):
        if kind == "rbf":
            X = torch.as_tensor(problems.rbf_points(n, seed=70 + r), device=device)
            mv, params = operators.RBF(X, precision="bf16x2"), (0.5, 1e-2)
        else:
            A = problems.dense_spd_device(n, seed=70 + r, shift=0.5, device=device)
            mv, params = operators.dense_matvec, (A,)
        reqs.append(serving.Request(mv, params, n, P, k, seed=r, arrival=0.002 * r,
                                    grad=(r % 3 == 2)))
    torch.cuda.synchronize()
    return reqs


def alone_full(req, device):
    Z = stochtrace.sampler_signs(np.zeros(req.n, np.float32), num=req.P).flat(
        prng.PRNGKey(req.seed), device)
    if req.grad:
        per, _, grads = funm.logdet_value_and_grad(req.matvec, Z, req.k, req.params)
        return per.cpu().numpy(), [g.cpu().numpy() for g in grads]
    per, _ = funm.logdet_per_probe(req.matvec, Z, req.k, req.params)
    return per.cpu().numpy(), None


def test_serving_distribution_mix_with_binding_budget(cuda):
    # the serving sizes with gradient requests, under a budget that fits the
    # largest request but not two large ones together: admissions wait, idle
    # slots give their scratch back, the accounted resident total stays within
    # the budget, and every result is bit-identical to the request alone
    reqs = c5_requests(cuda, np.random.default_rng(5))
    ref = [alone_full(q, cuda) for q in reqs]
    need = [serving.footprint(q, cuda) for q in reqs]
    budget = max(need) + sorted(need)[len(need) // 2]
    server = serving.Server(cuda, slots=4, hbm_budget=budget)
    out = server.run(reqs)
    st = server.stats
    print("serving stats", st, flush=True)
    assert st["waited_for_memory"] > 0 and st["releases"] > 0
    assert st["peak_resident"] <= budget
    assert sum(c.scratch_bytes() for c in server.contexts) <= budget
    overlap = any(a.start < b.finish and b.start < a.finish
                  for i, a in enumerate(out) for b in out[i + 1:])
    assert overlap
    for o, (per, grads) in zip(out, ref):
        assert o.error is None
        np.testing.assert_array_equal(o.per_probe.cpu().numpy(), per)
        if grads is not None:
            for g, gr in zip(o.grads, grads):
                np.testing.assert_array_equal(g.cpu().numpy(), gr)


def test_a_failing_request_does_not_stop_the_others(cuda):
    reqs = mixed_requests(cuda)[:3]
    bad = reqs[1]._replace(k=reqs[1].n + 1)  # num_matvecs > n: "exceeds"
    out = serving.Server(cuda, slots=2).run([reqs[0], bad, reqs[2]])
    assert out[1].error is not None and "exceeds" in out[1].error
    assert out[1].per_probe is None
    for o, q in zip((out[0], out[2]), (reqs[0], reqs[2])):
        assert o.error is None
        np.testing.assert_array_equal(o.per_probe.cpu().numpy(), alone(q, cuda))


def test_deadline_cancels_a_request_that_cannot_start(cuda):
    reqs = mixed_requests(cuda)
    big = reqs[4]  # n = 5000, 16 probes, depth 30
    late = reqs[5]._replace(arrival=0.0, deadline=0.0)
    server = serving.Server(cuda, slots=1)
    out = server.run([big, late])
    assert out[0].error is None
    assert out[1].error is not None and out[1].error.startswith("cancelled")
    assert server.stats["cancelled"] == 1
