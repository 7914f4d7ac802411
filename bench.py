#!/usr/bin/env python3
"""Throughput of the SLQ log-determinant path on one MI355X.

One step = one full SLQ log-determinant solve: Rademacher probes ->
tridiag_sym(k, reortho="full") -> log-det integrand -> probe mean, with all
inputs resident in HBM.  The metric counts probe-Lanczos-steps (one probe x one
Arnoldi step), so ``value = probes * depth * steps / seconds``.

Default workload: the matrix-free RBF log-determinant, n = 32768 points
uniform in [0, 1]^3 (seeded), K_ij = exp(-|x_i - x_j|^2 / (2 * 0.5^2)) +
1e-2 delta_ij, 128 probes, depth 64.  ``--workload dense`` runs the stored
dense SPD case (n = 16384, 256 probes, depth 64, A = B B^T / n + 0.1 I);
``rbf-none`` the RBF case with ``tridiag_sym(64, reortho="none")`` (plain
Lanczos, no basis); ``rbf-funm`` ``funm_lanczos_sym`` (log(A) v for the 128
probe vectors, depth 64); ``rbf-grad`` value and gradient at n = 8192, 32 probes,
depth 32; ``gp-logpdf`` the Gaussian log-density of the GP tutorial (SLQ
log-determinant + CG, value and gradient) at n = 8192.

Printed: one JSON line with the throughput, the relative error against an
FP64 dense-Cholesky log-determinant of the same operator, the roofline of the
dominant kernel (device time from HIP events on the launch stream during the
timed region) and a CPU baseline (NumPy restatement of matfree's algorithm on
a bounded sample of the same workload).
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SLQ probe-Lanczos-steps/s at ≤1e-3 rel err vs FP64; serve p50/p99 latency"
UNIT = "probe-Lanczos-steps/s"
PEAK = {  # MI355X published dense peaks
    "hbm_GBs": 8000.0,
    "fp32_mfma_TFs": 157.3,
    "bf16_mfma_TFs": 2500.0,
}
WORKLOADS = {
    "rbf": dict(n=32768, probes=128, depth=64, lengthscale=0.5, noise=1e-2),
    "dense": dict(n=16384, probes=256, depth=64, shift=0.1),
    "rbf-grad": dict(n=8192, probes=32, depth=32, lengthscale=0.5, noise=1e-2),
    "rbf-none": dict(n=32768, probes=128, depth=64, lengthscale=0.5, noise=1e-2),
    "rbf-stored": dict(n=32768, probes=128, depth=64, lengthscale=0.5, noise=1e-2),
    "rbf-funm": dict(n=32768, probes=128, depth=64, lengthscale=0.5, noise=1e-2),
    "gp-logpdf": dict(n=8192, probes=32, depth=32, lengthscale=0.5, noise=1e-2, rank=128),
}
DEFAULT_PRECISION = {"rbf": "bf16x2", "dense": "fp32", "rbf-grad": "fp32", "rbf-none": "bf16x2",
                     "rbf-funm": "bf16x2", "gp-logpdf": "fp32", "rbf-stored": "fp32"}
# packed storage of the stored matrix (dense, rbf-stored): fp32 = the plain
# FP32 matrix; bf16x2 / bf16 / fp8x2 / fp8 = operators.pack forms
DEFAULT_STORAGE = {"dense": "fp32", "rbf-stored": "fp8x2"}
# Serving mix: n log-uniform in [2048, 32768], probes in {16, 32, 64, 128},
# depth uniform in [16, 64], RBF (bf16x2) or stored dense (fp32, n capped at
# 16384 to bound the per-request matrix at 1 GB).
SERVE_N = (2048, 32768)
SERVE_DENSE_MAX_N = 16384
SERVE_PROBES = (16, 32, 64, 128)
SERVE_SLOTS = 4  # requests in flight at once in the serving stream
SERVE_DEPTH = (16, 64)
SERVE_GRAD_EVERY = 3  # every third request asks for the gradient too


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="rbf")
    ap.add_argument("--precision", default=None, help="fp32 | bf16 | bf16x2 | bf16x3")
    ap.add_argument("--storage", default=None,
                    help="stored matrix (dense, rbf-stored): fp32 | bf16x2 | bf16 | fp8x2 | fp8")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-reference", action="store_true", help="skip the FP64 Cholesky check")
    ap.add_argument("--serve-requests", type=int, default=48,
                    help="requests of the serving measurement (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=40.0,
                    help="arrival rate of the Poisson request stream (requests/s)")
    ap.add_argument("--serve-budget-gib", type=float, default=2.0,
                    help="HBM budget of the serving loop for the solves' scratch (GiB)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def setup(args, device):
    import numpy as np
    import torch

    from matfree_amd import operators, problems, prng, stochtrace

    w = WORKLOADS[args.workload]
    n, P, k = w["n"], w["probes"], w["depth"]
    prec = args.precision or DEFAULT_PRECISION[args.workload]
    storage = args.storage or DEFAULT_STORAGE.get(args.workload, "fp32")
    if args.workload in ("rbf", "rbf-none", "rbf-funm"):
        X = problems.rbf_points(n, seed=args.seed)
        op = operators.RBF(X, precision=prec)
        params = (w["lengthscale"], w["noise"])
        host = {"points": X}
    else:
        if args.workload == "rbf-stored":
            # the RBF kernel matrix of the points, stored (FP64 entries rounded once)
            X = problems.rbf_points(n, seed=args.seed)
            K = problems.rbf_dense64_device(X, w["lengthscale"], w["noise"], device)
            A = K.to(torch.float32)
            del K
            torch.cuda.empty_cache()
            host = {"points": X}
        else:
            A = problems.dense_spd_device(n, seed=args.seed, shift=w["shift"], device=device)
            host = {"matrix": A}
        if storage == "fp32":
            op = operators.DenseMatvec(precision=prec)
            params = (A,)
            host["matrix"] = A
        else:
            op = operators.packed_matvec
            params = (operators.pack(A, storage),)
            prec = "stored-" + storage
            if args.workload == "rbf-stored":
                del A
                torch.cuda.empty_cache()
    Z = stochtrace.sampler_signs(np.zeros(n, np.float32), num=P).flat(prng.PRNGKey(args.seed),
                                                                       device)
    torch.cuda.synchronize()
    return op, params, Z, (n, P, k), prec, host


def fp64_reference(args, host, device):
    import torch

    from matfree_amd import problems

    w = WORKLOADS[args.workload]
    if args.workload in ("rbf", "rbf-none", "rbf-stored"):
        K = problems.rbf_dense64_device(host["points"], w["lengthscale"], w["noise"], device)
    else:
        K = host["matrix"].to(torch.float64)
    val = problems.logdet_cholesky_device(K)
    del K
    torch.cuda.empty_cache()
    return val


def timed_solves(op, params, Z, k, steps, warmup, ctx, reortho="full"):
    import torch

    from matfree_amd import funm

    classes = ("matvec", "dots", "project_dots", "project", "normalize", "quadrature")
    for _ in range(warmup):
        funm.logdet_per_probe(op, Z, k, params, reortho=reortho)
    torch.cuda.synchronize()
    # timed region: events around the block matvec only (the roofline's
    # kernel); every event record costs device time
    ctx.profile(True, ("matvec",))
    for kc in classes:
        ctx.profile_read(kc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mean = None
    for _ in range(steps):
        _, mean = funm.logdet_per_probe(op, Z, k, params, reortho=reortho)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kernels = {"matvec": dict(zip(("total_ms", "launches"), ctx.profile_read("matvec")))}
    # per-class breakdown from as many further solves with every class bracketed
    ctx.profile(True)
    for _ in range(steps):
        funm.logdet_per_probe(op, Z, k, params, reortho=reortho)
    for kc in classes[1:]:
        ms, cnt = ctx.profile_read(kc)
        kernels[kc] = {"total_ms": ms, "launches": cnt}
    ctx.profile(False)
    return t1 - t0, float(mean), kernels


def roofline(kernels, n, P, k, prec, workload):
    """Roofline entry of the kernel class with the largest device time."""
    name = max(kernels, key=lambda kc: kernels[kc]["total_ms"])
    kk = kernels[name]
    avg_s = kk["total_ms"] / max(kk["launches"], 1) / 1e3
    if name == "matvec":
        flops = 2.0 * n * n * P  # algorithmic flops of one block matvec
        peak = PEAK["fp32_mfma_TFs"] if prec == "fp32" else PEAK["bf16_mfma_TFs"]
        achieved = flops / avg_s / 1e12
        entry = {"kernel": "block_matvec_" + workload + "_" + prec, "bound": "mfma",
                 "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                 "algorithmic_per_launch": flops, "avg_launch_ms": avg_s * 1e3}
        if prec.startswith("stored-"):
            # the stored planes streamed once per probe block
            per_el = {"bf16x2": 4, "bf16": 2, "fp8x2": 2, "fp8": 1}[prec[len("stored-"):]]
            blocks = -(-P // 128)
            entry["matrix_bytes_per_launch"] = float(per_el) * n * n * blocks
            entry["matrix_GBs"] = entry["matrix_bytes_per_launch"] / avg_s / 1e9
    else:
        # Streaming Gram-Schmidt kernels: bytes per launch averaged over the
        # k steps (column j reads j + 1 basis vectors of n floats per probe).
        mean_cols = (k + 1) / 2.0
        per_launch = {
            "dots": 4.0 * n * P * (mean_cols + 1),
            "project_dots": 4.0 * n * P * (mean_cols + 2),
            "project": 4.0 * n * P * (mean_cols + 2),
            "normalize": 8.0 * n * P,
            "quadrature": 8.0 * P * k * k,
        }[name]
        achieved = per_launch / avg_s / 1e9
        entry = {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": PEAK["hbm_GBs"],
                 "unit": "GB/s", "frac": achieved / PEAK["hbm_GBs"],
                 "algorithmic_per_launch": per_launch, "avg_launch_ms": avg_s * 1e3}
    entry["traffic"] = measured_traffic(entry["kernel"])
    return entry


def request_roofline_ms(kind, n, P, k, grad):
    """Lower bound of one serving request's device time (ms) at the roofline:
    its block-matvec work (2 n^2 P algorithmic flops per matvec, k matvecs;
    a gradient request adds the reverse pass's k matvecs and the parameter
    contraction, 2 n^2 P k flops) at the matrix-core peak of its precision
    (BF16 for the RBF bf16x2 requests, FP32 for the stored FP32 ones), plus
    the Gram-Schmidt streams at the HBM peak: forward step j reads the j + 1
    basis columns three times and moves ~5 more P x n vectors; a reverse step
    idx reads ~2 (idx + 2) + (idx + 1) + (2 k - idx - 1) columns and ~8 vectors
    (dots, projection, the second dots, the adjoint combination)."""
    peak = PEAK["bf16_mfma_TFs"] if kind == "rbf" else PEAK["fp32_mfma_TFs"]
    vec = 4.0 * P * n
    mv = 2.0 * n * n * P * k * (3 if grad else 1)
    gs = vec * (3.0 * k * (k + 1) / 2.0 + 5.0 * k)
    if grad:
        gs += vec * sum(2 * (i + 2) + (i + 1) + (2 * k - i - 1) + 8 for i in range(k))
    return (mv / (peak * 1e12) + gs / (PEAK["hbm_GBs"] * 1e9)) * 1e3


def measured_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
        return table.get(kernel, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def serve(args, device):
    """Latency of a stream of independent SLQ log-determinant (+ gradient) solves.

    Each request draws its operator (RBF on fresh points, or a stored SPD
    matrix), dimension, probe count and depth from the serving mix, and every
    third request asks for the gradient too (value and d/d parameters through
    the reverse pass); all operators are made resident in HBM first.  Closed
    loop: every request (probe generation, solve, result ready on the device)
    is timed alone.  Then the same requests arrive as a Poisson stream at a
    fixed rate (requests/s) and run through the continuous-batching loop of
    matfree_amd.serving (4 slots, each a stream with its own library context,
    admitted in arrival order within an HBM budget for the solves' scratch):
    latency = completion - arrival.  Each request's per-probe results (and
    gradients) in the mix are compared with its results alone (bit for bit
    expected), and the same arrivals are replayed through one FIFO server with
    the closed-loop service times for comparison.
    """
    import numpy as np
    import torch

    from matfree_amd import funm, operators, problems, prng, serving, stochtrace

    rng = np.random.default_rng(args.seed + 1)
    reqs, meta = [], []
    for r in range(args.serve_requests):
        n = int(round(np.exp(rng.uniform(np.log(SERVE_N[0]), np.log(SERVE_N[1])))))
        P = int(rng.choice(SERVE_PROBES))
        k = int(rng.integers(SERVE_DEPTH[0], SERVE_DEPTH[1] + 1))
        kind = "rbf" if (rng.random() < 0.5 or n > SERVE_DENSE_MAX_N) else "dense"
        if kind == "rbf":
            X = torch.as_tensor(problems.rbf_points(n, seed=1000 + r), device=device)
            op, params = operators.RBF(X, precision="bf16x2"), (0.5, 1e-2)
        else:
            A = problems.dense_spd_device(n, seed=1000 + r, shift=0.1, device=device)
            op, params = operators.DenseMatvec(precision="fp32"), (A,)
        grad = r % SERVE_GRAD_EVERY == SERVE_GRAD_EVERY - 1
        reqs.append(serving.Request(op, params, n, P, k, seed=r, grad=grad))
        meta.append((kind, n, P, k, grad))
    # start-up: one small untimed value and value+gradient solve per operator
    # kind, as a server does before taking traffic (loads the library and its
    # kernels, creates the context); every request below is then timed
    wn, wP, wk = 1000, 16, 8
    wX = torch.as_tensor(problems.rbf_points(wn, seed=999), device=device)
    wA = problems.dense_spd_device(wn, seed=999, shift=0.1, device=device)
    wZ = torch.ones((wP, wn), device=device)
    for wop, wpar in ((operators.RBF(wX, precision="bf16x2"), (0.5, 1e-2)),
                      (operators.DenseMatvec(precision="fp32"), (wA,))):
        funm.logdet_per_probe(wop, wZ, wk, wpar)
        funm.logdet_value_and_grad(wop, wZ, wk, wpar)
    del wX, wA, wZ
    torch.cuda.synchronize()
    lat, alone = [], []
    for q in reqs:
        sampler = stochtrace.sampler_signs(np.zeros(q.n, np.float32), num=q.P)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Z = sampler.flat(prng.PRNGKey(q.seed), device)
        grads = None
        if q.grad:
            per, mean, grads = funm.logdet_value_and_grad(q.matvec, Z, q.k, q.params)
        else:
            per, mean = funm.logdet_per_probe(q.matvec, Z, q.k, q.params)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t0) * 1e3)
        alone.append((per.cpu().numpy(), None if grads is None else
                      [g.cpu().numpy() for g in grads]))
        del Z, per, mean, grads
    lat = np.array(lat)
    # Poisson arrivals at a fixed rate
    rate = args.serve_rate
    arrive = np.cumsum(rng.exponential(1.0 / rate, size=len(lat)))  # seconds
    mixed = [q._replace(arrival=float(a)) for q, a in zip(reqs, arrive)]
    budget = int(args.serve_budget_gib * 2**30)
    torch.cuda.synchronize()
    server = serving.Server(device, slots=SERVE_SLOTS, hbm_budget=budget)
    t0 = time.perf_counter()
    out = server.run(mixed)
    span = time.perf_counter() - t0
    errors = [o.error for o in out if o.error is not None]
    resp = np.array([(o.finish - o.arrival) * 1e3 for o in out])

    def rdiff(a, b):
        a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
        return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))

    diff = 0.0
    for o, (per, grads) in zip(out, alone):
        if o.error is None:
            diff = max(diff, rdiff(o.per_probe.cpu().numpy(), per))
            if grads is not None:
                for g, gr in zip(o.grads, grads):
                    diff = max(diff, rdiff(g.cpu().numpy(), gr))
    # the same arrivals through one FIFO server with the service times alone
    finish = np.empty_like(arrive)
    busy = 0.0
    for i in range(len(lat)):
        start = max(arrive[i] * 1e3, busy)
        finish[i] = busy = start + lat[i]
    fifo = finish - arrive * 1e3
    st = server.stats
    bound_ms = sum(request_roofline_ms(*m) for m in meta)
    val = np.array([t for t, m in zip(lat, meta) if not m[4]])
    grd = np.array([t for t, m in zip(lat, meta) if m[4]])
    del reqs, mixed, out, server
    torch.cuda.empty_cache()
    return {
        "requests": len(lat),
        "gradient_requests": int(sum(m[4] for m in meta)),
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()),
        "max_ms": float(lat.max()),
        "value_only": {"requests": int(len(val)), "p50_ms": float(np.percentile(val, 50)),
                       "p99_ms": float(np.percentile(val, 99))},
        "with_gradient": {"requests": int(len(grd)), "p50_ms": float(np.percentile(grd, 50)),
                          "p99_ms": float(np.percentile(grd, 99))},
        # the mix's work at the roofline (request_roofline_ms) against the
        # requests' summed latencies alone and the stream's wall time
        "roofline": {"lower_bound_ms": bound_ms, "frac_alone": bound_ms / float(lat.sum()),
                     "frac_stream": bound_ms / (span * 1e3)},
        "stream": {
            "arrival_rate_per_s": rate,
            "slots": SERVE_SLOTS,
            "hbm_budget_bytes": budget,
            "p50_ms": float(np.percentile(resp, 50)),
            "p99_ms": float(np.percentile(resp, 99)),
            "requests_per_s": len(lat) / span,
            "probe_steps_per_s": sum(m[2] * m[3] for m in meta) / span,
            "max_rel_diff_vs_alone": diff,
            "errors": len(errors),
            "admissions_waited_for_memory": st.get("waited_for_memory"),
            "scratch_releases": st.get("releases"),
            "peak_resident_bytes": st.get("peak_resident"),
            "fifo_one_at_a_time_p50_ms": float(np.percentile(fifo, 50)),
            "fifo_one_at_a_time_p99_ms": float(np.percentile(fifo, 99)),
        },
        "mix": (f"n log-uniform [{SERVE_N[0]}, {SERVE_N[1]}], probes {list(SERVE_PROBES)}, "
                f"depth U[{SERVE_DEPTH[0]}, {SERVE_DEPTH[1]}], RBF bf16x2 or dense fp32 "
                f"(dense n <= {SERVE_DENSE_MAX_N}); {sum(m[0] == 'rbf' for m in meta)} RBF; "
                f"every {SERVE_GRAD_EVERY}rd request value + gradient"),
    }


def grad_step(args, device):
    """Value and gradient (d/d lengthscale, d/d noise) of the RBF estimate."""
    import numpy as np
    import torch

    from matfree_amd import _runtime, funm, operators, problems, prng, stochtrace

    w = WORKLOADS["rbf-grad"]
    n, P, k = w["n"], w["probes"], w["depth"]
    prec = args.precision or DEFAULT_PRECISION["rbf-grad"]
    X = torch.as_tensor(problems.rbf_points(n, seed=args.seed), device=device)
    op = operators.RBF(X, precision=prec)
    Z = stochtrace.sampler_signs(np.zeros(n, np.float32), num=P).flat(prng.PRNGKey(args.seed),
                                                                       device)
    params = (w["lengthscale"], w["noise"])
    for _ in range(args.warmup):
        funm.logdet_value_and_grad(op, Z, k, params)
    ctx = _runtime.context(device)
    ctx.profile(True, ("adjoint",))
    for kc in ("matvec", "adjoint"):
        ctx.profile_read(kc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, mean, grads = funm.logdet_value_and_grad(op, Z, k, params)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    adj_ms, adj_n = ctx.profile_read("adjoint")
    ctx.profile(False)
    _, fmean = funm.logdet_per_probe(op, Z, k, params)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        funm.logdet_per_probe(op, Z, k, params)
    torch.cuda.synchronize()
    fwd = (time.perf_counter() - t1) / args.steps
    return {
        "metric": "SLQ log-det value + gradient (d/d lengthscale, d/d noise) time",
        "value": dt * 1e3,
        "unit": "ms",
        "higher_is_better": False,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "dtype": prec,
        "data": "synthetic (seeded points, Rademacher probes from PRNGKey(seed))",
        "config": {"workload": f"rbf-grad n={n} probes={P} depth={k}", "n": n, "probes": P,
                   "depth": k, "precision": prec, "lengthscale": w["lengthscale"],
                   "noise": w["noise"]},
        "forward_ms": fwd * 1e3,
        "reverse_pass_ms": adj_ms / max(adj_n, 1),
        "logdet": float(mean),
        "grad": [float(grads[0]), float(grads[1])],
    }


def funm_step(args, device):
    """funm_lanczos_sym(dense_funm_sym_eigh(log), tridiag_sym(64)): log(A) v for the
    128 probe vectors of the RBF problem, with the FP32 form beside it."""
    import numpy as np
    import torch

    from matfree_amd import _lib, _runtime, funm, operators

    op, params, Z, (n, P, k), prec, _ = setup(args, device)
    code = _lib.MATFUNS["log"]
    ctx = _runtime.context(device)
    for _ in range(args.warmup):
        funm.funm_vectors(op, Z, k, params, code)
    torch.cuda.synchronize()
    classes = ("matvec", "dots", "project_dots", "project", "normalize")
    ctx.profile(True, ("matvec",))  # timed region: the roofline kernel's events only
    for kc in classes:
        ctx.profile_read(kc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        Y = funm.funm_vectors(op, Z, k, params, code)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    kernels = {"matvec": dict(zip(("total_ms", "launches"), ctx.profile_read("matvec")))}
    ctx.profile(True)  # per-class breakdown from further calls
    for _ in range(args.steps):
        funm.funm_vectors(op, Z, k, params, code)
    for kc in classes[1:]:
        ms, cnt = ctx.profile_read(kc)
        kernels[kc] = {"total_ms": ms, "launches": cnt}
    ctx.profile(False)
    out = {
        "metric": "funm_lanczos_sym probe-Lanczos-steps/s (log(A) v)",
        "value": P * k / dt,
        "unit": UNIT,
        "higher_is_better": True,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "dtype": prec + "+fp32",
        "data": "synthetic (seeded points, Rademacher vectors from PRNGKey(seed))",
        "config": {"workload": f"rbf-funm n={n} vectors={P} depth={k}", "n": n, "probes": P,
                   "depth": k, "precision": prec, "matfun": "log",
                   "lengthscale": WORKLOADS["rbf-funm"]["lengthscale"],
                   "noise": WORKLOADS["rbf-funm"]["noise"]},
        "roofline": roofline(kernels, n, P, k, prec, "rbf"),
        "kernels_ms_per_solve": {kc: v["total_ms"] / args.steps for kc, v in kernels.items()},
    }
    if not args.no_reference:
        # the same f(A) V with the FP32 matvec: the BF16 form's deviation
        op32 = operators.RBF(op.points, precision="fp32")
        Y32 = funm.funm_vectors(op32, Z, k, params, code)
        out["rel_diff_vs_fp32_form"] = float(torch.linalg.norm(Y - Y32) / torch.linalg.norm(Y32))
    return out


def gp_step(args, device):
    """The GP tutorial's log-density, value and gradient (tutorials/8_gaussian_logpdf.py).

    logpdf(y) = -(y . A^{-1} y + logdet A + n log(2 pi)) / 2 with the SLQ
    estimate of logdet A (Arnoldi, FP32) and CG for A^{-1} y, differentiated
    in (lengthscale, noise) by torch autograd; beside it CG with the rank-r
    pivoted partial-Cholesky preconditioner of K, and the dense FP64 value.
    """
    import math

    import numpy as np
    import torch

    from matfree_amd import _runtime, gp, linalg, low_rank, operators, problems, prng

    w = WORKLOADS["gp-logpdf"]
    n, P, k, rank = w["n"], w["probes"], w["depth"], w["rank"]
    prec = args.precision or DEFAULT_PRECISION["gp-logpdf"]
    Xh = problems.rbf_points(n, seed=args.seed)
    X = torch.as_tensor(Xh, device=device)
    op = operators.RBF(X, precision=prec)
    y = torch.as_tensor(np.random.default_rng(args.seed).standard_normal(n).astype(np.float32),
                        device=device)
    logpdf = gp.create_logpdf(op, np.zeros(n, np.float32), num_matvecs=k, num_samples=P)
    key = prng.PRNGKey(args.seed)

    def value_and_grad():
        ls = torch.tensor(w["lengthscale"], dtype=torch.float64, device=device, requires_grad=True)
        s2 = torch.tensor(w["noise"], dtype=torch.float64, device=device, requires_grad=True)
        val = logpdf(y, key, ls, s2)
        val.backward()
        return val.detach(), ls.grad, s2.grad

    for _ in range(args.warmup):
        value_and_grad()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        val, gl, gs = value_and_grad()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    params = (w["lengthscale"], w["noise"])
    ctx = _runtime.context(device)
    ctx.profile(True, ("matvec",))
    ctx.profile_read("matvec")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    x0, it0 = linalg.cg_batched(op, y, params)
    torch.cuda.synchronize()
    t_cg = time.perf_counter() - t1
    mv_ms, mv_n = ctx.profile_read("matvec")
    ctx.profile(False)
    t2 = time.perf_counter()
    M = low_rank.preconditioner(low_rank.cholesky_partial_pivot(op, nrows=n, rank=rank)).bind(
        w["noise"], w["lengthscale"], 0.0)
    torch.cuda.synchronize()
    t_chol = time.perf_counter() - t2
    t3 = time.perf_counter()
    x1, it1 = linalg.cg_batched(op, y, params, M=M)
    torch.cuda.synchronize()
    t_pcg = time.perf_counter() - t3
    out = {
        "metric": "GP log-density value + gradient (d/d lengthscale, d/d noise) time",
        "value": dt * 1e3,
        "unit": "ms",
        "higher_is_better": False,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "dtype": prec,
        "data": "synthetic (seeded points and targets, Rademacher probes from PRNGKey(seed))",
        "config": {"workload": f"gp-logpdf n={n} probes={P} depth={k}", "n": n, "probes": P,
                   "depth": k, "precision": prec, "lengthscale": w["lengthscale"],
                   "noise": w["noise"], "cg_tol": 1e-5},
        "logpdf": float(val),
        "grad": [float(gl), float(gs)],
        "cg": {"iterations": int(it0), "ms": t_cg * 1e3,
               "matvec_ms_per_iteration": mv_ms / max(mv_n, 1)},
        "pcg": {"rank": rank, "iterations": int(it1), "ms": t_pcg * 1e3,
                "cholesky_ms": t_chol * 1e3},
    }
    if not args.no_reference:
        K = problems.rbf_dense64_device(Xh, w["lengthscale"], w["noise"], device)
        L = torch.linalg.cholesky(K)
        y64 = y.to(torch.float64)
        alpha = torch.cholesky_solve(y64[:, None], L)[:, 0]
        logdet = 2.0 * float(torch.log(torch.diagonal(L)).sum())
        ref = -0.5 * (float(y64 @ alpha) + logdet + n * math.log(2 * math.pi))
        out["logpdf_fp64"] = ref
        out["rel_err_vs_fp64"] = abs(float(val) - ref) / abs(ref)
        out["cg_rel_err_vs_fp64"] = float(torch.linalg.norm(x0.to(torch.float64) - alpha)
                                          / torch.linalg.norm(alpha))
        out["pcg_rel_err_vs_fp64"] = float(torch.linalg.norm(x1.to(torch.float64) - alpha)
                                           / torch.linalg.norm(alpha))
        del K, L
        torch.cuda.empty_cache()
    return out


def cpu_baseline(args, host, device, n, k):
    """NumPy restatement of matfree's algorithm (the oracle) on a bounded sample."""
    import numpy as np
    import torch

    from oracle import prng as oprng
    from oracle import slq as oslq

    w = WORKLOADS[args.workload]
    probes = 96  # about 10-20 s of host work at the restatement's rate
    if args.workload in ("rbf", "rbf-none", "rbf-stored"):
        from matfree_amd import problems

        K = problems.rbf_dense64_device(host["points"], w["lengthscale"], w["noise"], device)
        A32 = K.to(torch.float32).cpu().numpy()
        del K
        torch.cuda.empty_cache()
    else:
        A32 = host["matrix"].cpu().numpy()
    op = oslq.DenseOperator(A32, np.float32)
    Z = oprng.rademacher(oprng.prng_key(args.seed), (probes, n), np.float32)
    try:
        from threadpoolctl import threadpool_info

        cores = max(int(i.get("num_threads", 1)) for i in threadpool_info()) if threadpool_info() else 1
    except Exception:
        cores = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    if args.workload == "rbf-none":
        oslq.quadform_logdet_lanczos(op, k, Z, np.float32)
        form = "batched A@V, three-term recurrence, eigh"
    else:
        oslq.quadform_logdet(op, k, Z, np.float32)
        form = "batched A@V, full-width CGS2, eigh"
    dt = time.perf_counter() - t0
    sample = (f"{probes} probes x {k} steps of the same n={n} {args.workload} operator, "
              f"FP32 NumPy restatement of matfree's formulation ({form}); operator "
              "materialised on the host before timing")
    return {"value": probes * k / dt, "unit": UNIT, "cores": cores, "kind": "port",
            "sample": sample, "seconds": dt}


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "runs on one GPU only: probes would shard across GPUs with one "
                                   "all-reduce of the per-probe values, which is not implemented",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    import torch

    from matfree_amd import _runtime

    device = torch.device("cuda", 0)
    torch.cuda.set_device(device)
    if args.workload == "rbf-grad":
        print(json.dumps(grad_step(args, device)), flush=True)
        return
    if args.workload == "gp-logpdf":
        print(json.dumps(gp_step(args, device)), flush=True)
        return
    if args.workload == "rbf-funm":
        print(json.dumps(funm_step(args, device)), flush=True)
        return
    op, params, Z, (n, P, k), prec, host = setup(args, device)
    ctx = _runtime.context(device)
    log(f"bench: workload={args.workload} n={n} P={P} k={k} precision={prec}")
    reortho = "none" if args.workload == "rbf-none" else "full"
    dt, est, kernels = timed_solves(op, params, Z, k, args.steps, args.warmup, ctx, reortho)
    value = P * k * args.steps / dt
    log(f"bench: {dt / args.steps * 1e3:.2f} ms/solve, {value:.0f} {UNIT}, logdet {est:.8e}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if prec == "fp32" else prec + "+fp32",
        "data": "synthetic (seeded points / matrix, Rademacher probes from PRNGKey(seed))",
        "config": {"workload": f"{args.workload}-logdet n={n} probes={P} depth={k}",
                   "n": n, "probes": P, "depth": k, "precision": prec,
                   **{kk: v for kk, v in WORKLOADS[args.workload].items()
                      if kk not in ("n", "probes", "depth")}},
        "logdet": est,
    }
    if not args.no_reference:
        ref = fp64_reference(args, host, device)
        out["logdet_fp64_cholesky"] = ref
        out["rel_err_vs_fp64"] = abs(est - ref) / abs(ref)
        log(f"bench: FP64 Cholesky {ref:.8e}, rel err {out['rel_err_vs_fp64']:.2e}")
    out["roofline"] = roofline(kernels, n, P, k, prec, args.workload)
    out["kernels_ms_per_solve"] = {kc: v["total_ms"] / args.steps for kc, v in kernels.items()}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args, host, device, n, k)
        log(f"bench: cpu baseline {out['cpu_baseline']['value']:.1f} {UNIT}")
    if args.serve_requests > 0:
        del op, params, Z, host
        torch.cuda.empty_cache()
        out["serve"] = serve(args, device)
        sv = out["serve"]
        log(f"bench: serve alone value-only p50 {sv['value_only']['p50_ms']:.2f} ms, p99 "
            f"{sv['value_only']['p99_ms']:.2f} ms; with gradient p50 "
            f"{sv['with_gradient']['p50_ms']:.2f} ms, p99 {sv['with_gradient']['p99_ms']:.2f} ms")
        log(f"bench: serve alone p50 {sv['p50_ms']:.2f} ms, p99 {sv['p99_ms']:.2f} ms; stream "
            f"p50 {sv['stream']['p50_ms']:.2f} ms, p99 {sv['stream']['p99_ms']:.2f} ms; "
            f"roofline fraction alone {sv['roofline']['frac_alone']:.3f}, stream "
            f"{sv['roofline']['frac_stream']:.3f}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
