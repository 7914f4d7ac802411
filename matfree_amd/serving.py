"""Continuous-batching request loop for SLQ log-determinant solves on one GPU.

A request is one SLQ solve -- the log-determinant estimate, or the estimate
and its parameter gradient -- of an operator already resident in HBM: its
parameters, a probe count P and key, and a depth k.  Up to ``slots`` requests
run at the same time, each on a stream of its own (default priority) with a
library context of its own, so that their kernels overlap on the device while
every request computes exactly what it computes alone (the kernels reduce in
fixed orders, and no scratch is shared).

Admission by HBM footprint.  A request's footprint is the device scratch its
library call needs (``slq_workspace_bytes``: the Krylov basis of k columns,
the work vectors, split-K partials and probe planes; plus its probes and
outputs), known exactly before it starts.  A slot's context keeps each
scratch buffer at the largest size any request it ran needed (the pool grows,
it does not shrink), so the loop accounts for what every slot holds:
resident = sum over slots of the scratch held (for a running request, what
the slot held plus the growth of its pool in this call, which the library
also computes exactly before the call: ``slq_workspace_growth``; a pool can
end above both its old holding and the request's fresh need when the two
differ in which buffers are large).  A request is admitted in arrival order once it has
arrived, a slot is free and the resident total with it fits the budget; if it
does not fit, the idle slots' scratch is released first (largest first); if
it still does not fit, the request waits for running ones to finish.  A
request larger than the whole budget runs when nothing else runs.  The
default budget is half of the device memory free when the server starts.

The loop runs on the host thread: it launches admitted requests (which
returns at once) and polls the completion events of the running ones.
Latency of a request = completion seen by the host - arrival time.  A
request that raises when it is launched (a bad argument, an allocation
failure) ends with its error and the others continue; a request whose
``deadline`` passes before it starts is cancelled.
"""

import collections
import time
from typing import Any, NamedTuple, Optional

import numpy as np
import torch

from . import _lib, funm, operators, prng, stochtrace


class Request(NamedTuple):
    """One solve: ``matvec(v, *params)`` of dimension n, P probes, depth k."""

    matvec: Any
    params: tuple
    n: int
    P: int
    k: int
    seed: int
    arrival: float = 0.0  # seconds after the start of Server.run
    grad: bool = False  # value and gradient (the reverse pass) instead of the value
    deadline: Optional[float] = None  # seconds after arrival by which it must start


class Result(NamedTuple):
    per_probe: Any  # (P,) float64 tensor on the device (None on error)
    mean: float
    arrival: float
    start: float
    finish: float  # completion seen by the host (seconds after the start)
    slot: int
    grads: Any = None  # parameter cotangents of the mean (grad requests)
    error: Optional[str] = None
    waited_for_memory: bool = False


def footprint(req: Request, device=None, context=None) -> int:
    """Device bytes a request needs beyond its operator: the library's scratch
    for the call (exact) plus its probes and outputs.  With ``context`` (a
    library context): the growth of that context's pooled scratch in the call,
    given what it holds, in place of a fresh context's scratch."""
    device = torch.device("cuda", torch.cuda.current_device()) if device is None else device
    op = operators.resolve(req.matvec)
    desc, keep = op.descriptor(req.n, req.params, device)
    entry = "logdet_vjp" if req.grad else "logdet"
    if context is None:
        ctx = _lib.context(device.index, torch.cuda.current_stream(device).cuda_stream)
        scratch = ctx.workspace_bytes(entry, desc, req.P, req.k)
    else:
        scratch = context.workspace_growth(entry, desc, req.P, req.k)
    del keep
    extra = 4 * req.P * req.n + 8 * req.P + 64
    if req.grad and not isinstance(op, operators.RBF):
        extra += 4 * req.n * req.n  # the matrix cotangent
    return int(scratch + extra)


def admission(need, slot, held, free, running_bytes, any_running, budget, growth=None):
    """Whether a request needing ``need`` bytes may start in idle ``slot``.

    ``held[s]``: scratch bytes slot s's context holds; ``free``: the idle
    slots; ``running_bytes``: bytes accounted to running requests;
    ``growth``: the bytes by which the slot's pool grows in the request (its
    context's ``footprint``); None for the bound ``need`` (a pool never grows
    by more than a fresh context needs).  The request is charged what the
    slot holds plus the growth, or ``need`` once the slot holds nothing.
    Returns ``(admit, release)``: admit or wait, and the idle slots whose
    scratch is to be released first (largest first, only as many as needed;
    none when the request waits).  A request that does not fit even with every
    idle slot released waits while anything runs, and runs alone otherwise.
    """
    if budget is None:
        return True, []
    # per-slot holdings, from a list indexed by slot or a mapping slot -> bytes
    h = dict(held) if isinstance(held, dict) else dict(enumerate(held))
    g = need if growth is None else growth

    def total():
        return sum(h[s] for s in free) + running_bytes + (g if h[slot] > 0 else need)

    release = []
    with_it = total()
    for s in sorted(free, key=lambda s: -h[s]):
        if with_it <= budget or h[s] == 0:
            break
        if s == slot and h[s] + g <= need:
            continue  # the request reuses this memory: releasing it gains nothing
        release.append(s)
        h[s] = 0
        with_it = total()
    if with_it <= budget or not any_running:
        return True, release
    return False, []


def default_budget(device) -> int:
    """Half of the device memory free now."""
    free, _ = torch.cuda.mem_get_info(device)
    return int(free // 2)


class Server:
    """Runs requests with up to ``slots`` in flight under an HBM budget.

    ``hbm_budget``: bytes, ``"auto"`` (:func:`default_budget`) or None (no
    limit)."""

    def __init__(self, device, slots: int = 4, hbm_budget="auto"):
        if slots < 1:
            raise ValueError("slots must be at least 1")
        self.device = torch.device(device)
        self.slots = int(slots)
        self.streams = [torch.cuda.Stream(self.device) for _ in range(self.slots)]
        self.contexts = [_lib.context(self.device.index, s.cuda_stream) for s in self.streams]
        self.hbm_budget = default_budget(self.device) if hbm_budget == "auto" else hbm_budget
        self.stats = {}

    def _held(self, slot):
        return self.contexts[slot].scratch_bytes()

    def _launch(self, req: Request, stream):
        stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(stream):
            sampler = stochtrace.sampler_signs(np.zeros(req.n, np.float32), num=req.P)
            Z = sampler.flat(prng.PRNGKey(req.seed), self.device)
            grads = None
            if req.grad:
                per, mean, grads = funm.logdet_value_and_grad(req.matvec, Z, req.k, req.params)
            else:
                per, mean = funm.logdet_per_probe(req.matvec, Z, req.k, req.params)
            done = torch.cuda.Event()
            done.record(stream)
        return per, mean, grads, done

    def run(self, requests, poll_sleep: float = 0.0):
        """Serve ``requests`` (arrival times relative to the call); returns one
        Result per request, in the order given.  ``self.stats`` afterwards holds
        the admission counters and the largest resident total accounted."""
        budget = self.hbm_budget
        need, bad = [], {}
        for i, q in enumerate(requests):
            try:
                need.append(footprint(q, self.device))
            except (ValueError, TypeError, NotImplementedError) as exc:
                need.append(0)
                bad[i] = f"{type(exc).__name__}: {exc}"
        order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
        queue = collections.deque(order)
        free = list(range(self.slots))
        running = {}  # slot -> (index, per, mean, grads, event, start, bytes accounted)
        results = [None] * len(requests)
        held = [self._held(s) for s in range(self.slots)]  # idle slots: scratch held
        waited = set()
        stats = {"waited_for_memory": 0, "releases": 0, "cancelled": 0, "failed": 0,
                 "peak_resident": sum(held), "budget": budget}

        def resident():
            return sum(held[s] for s in free) + sum(r[6] for r in running.values())

        t0 = time.perf_counter()
        while queue or running:
            now = time.perf_counter() - t0
            for slot in list(running):
                i, per, mean, grads, ev, start, _ = running[slot]
                if ev.query():
                    now = time.perf_counter() - t0
                    results[i] = Result(per, float(mean), requests[i].arrival, start, now, slot,
                                        grads, None, i in waited)
                    del running[slot]
                    free.append(slot)
                    held[slot] = self._held(slot)
            while queue and free and requests[queue[0]].arrival <= now:
                i = queue[0]
                q = requests[i]
                if i in bad:  # rejected by the library's argument checks
                    queue.popleft()
                    stats["failed"] += 1
                    results[i] = Result(None, float("nan"), q.arrival, now, now, -1, None,
                                        bad[i], False)
                    continue
                if q.deadline is not None and now > q.arrival + q.deadline:
                    queue.popleft()
                    stats["cancelled"] += 1
                    results[i] = Result(None, float("nan"), q.arrival, now, now, -1, None,
                                        f"cancelled: not started within {q.deadline} s", False)
                    continue
                slot = free[0]
                running_bytes = sum(r[6] for r in running.values())
                grow = (footprint(q, self.device, self.contexts[slot])
                        if budget is not None and held[slot] > 0 else need[i])
                admit, release = admission(need[i], slot, held, free, running_bytes,
                                           bool(running), budget, grow)
                if not admit:
                    if i not in waited:
                        waited.add(i)
                        stats["waited_for_memory"] += 1
                    break
                for s in release:  # give idle slots' memory back
                    self.contexts[s].release()
                    stats["releases"] += 1
                    held[s] = 0
                queue.popleft()
                free.pop(0)
                try:
                    per, mean, grads, ev = self._launch(q, self.streams[slot])
                except (ValueError, TypeError, MemoryError, NotImplementedError,
                        _lib.SLQError) as exc:
                    stats["failed"] += 1
                    results[i] = Result(None, float("nan"), q.arrival, now, now, slot, None,
                                        f"{type(exc).__name__}: {exc}", i in waited)
                    free.insert(0, slot)
                    held[slot] = self._held(slot)
                    continue
                running[slot] = (i, per, mean, grads, ev, now,
                                 held[slot] + grow if held[slot] > 0 else need[i])
                stats["peak_resident"] = max(stats["peak_resident"], resident())
            if poll_sleep > 0.0:
                time.sleep(poll_sleep)
        self.stats = stats
        return results
