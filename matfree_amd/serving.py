"""Continuous-batching request loop for SLQ log-determinant solves on one GPU.

A request is one ``funm.logdet_per_probe`` solve: an operator already
resident in HBM, its parameters, a probe count P and key, and a depth k.
Up to ``slots`` requests run at the same time, each on a stream of its own
(default priority) with a library context of its own, so that their kernels
overlap on the device while every request computes exactly what it computes
alone (the kernels reduce in fixed orders, and no scratch is shared).  A
request is admitted in arrival order once it has arrived, a slot is free and
its HBM footprint fits beside the footprints of the requests running:
``4 (k + 24) P n`` bytes (the FP32 basis of k columns, about 24 further
vectors of P x n floats for the Gram-Schmidt work vectors, probe planes and
split-K partials); operators are not counted (their memory is the caller's).

The loop runs on the host thread: it launches admitted requests (which
returns at once), and polls the completion events of the running ones.
Latency of a request = completion seen by the host - arrival time.
"""

import collections
import time
from typing import Any, NamedTuple, Optional

import numpy as np
import torch

from . import funm, prng, stochtrace


class Request(NamedTuple):
    """One solve: ``matvec(v, *params)`` of dimension n, P probes, depth k."""

    matvec: Any
    params: tuple
    n: int
    P: int
    k: int
    seed: int
    arrival: float = 0.0  # seconds after the start of Server.run


class Result(NamedTuple):
    per_probe: Any  # (P,) float64 tensor on the device
    mean: float
    arrival: float
    start: float
    finish: float  # completion seen by the host (seconds after the start)
    slot: int


def footprint(req: Request) -> int:
    """HBM bytes a request needs beyond its operator (its admission weight)."""
    return 4 * (req.k + 24) * req.P * req.n


def admits(req: Request, in_use: int, budget: Optional[int]) -> bool:
    """Whether ``req`` may start beside running requests of ``in_use`` bytes.

    A request larger than the whole budget is admitted when nothing else
    runs, so that it cannot wait forever."""
    if budget is None or in_use == 0:
        return True
    return in_use + footprint(req) <= budget


class Server:
    """Runs requests with up to ``slots`` in flight under an HBM budget."""

    def __init__(self, device, slots: int = 4, hbm_budget: Optional[int] = None):
        if slots < 1:
            raise ValueError("slots must be at least 1")
        self.device = device
        self.slots = int(slots)
        self.streams = [torch.cuda.Stream(device) for _ in range(self.slots)]
        self.hbm_budget = hbm_budget

    def _launch(self, req: Request, stream):
        stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(stream):
            sampler = stochtrace.sampler_signs(np.zeros(req.n, np.float32), num=req.P)
            Z = sampler.flat(prng.PRNGKey(req.seed), self.device)
            per, mean = funm.logdet_per_probe(req.matvec, Z, req.k, req.params)
            done = torch.cuda.Event()
            done.record(stream)
        return per, mean, done

    def run(self, requests, poll_sleep: float = 0.0):
        """Serve ``requests`` (arrival times relative to the call); returns one
        Result per request, in the order given."""
        order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
        queue = collections.deque(order)
        free = list(range(self.slots))
        running = {}  # slot -> (index, per, mean, event, start, bytes)
        results = [None] * len(requests)
        in_use = 0
        t0 = time.perf_counter()
        while queue or running:
            now = time.perf_counter() - t0
            for slot in list(running):
                i, per, mean, ev, start, fp = running[slot]
                if ev.query():
                    now = time.perf_counter() - t0
                    results[i] = Result(per, float(mean), requests[i].arrival, start, now, slot)
                    del running[slot]
                    free.append(slot)
                    in_use -= fp
            while (queue and free and requests[queue[0]].arrival <= now
                   and admits(requests[queue[0]], in_use, self.hbm_budget)):
                i = queue.popleft()
                slot = free.pop(0)
                per, mean, ev = self._launch(requests[i], self.streams[slot])
                fp = footprint(requests[i])
                in_use += fp
                running[slot] = (i, per, mean, ev, now, fp)
            if poll_sleep > 0.0:
                time.sleep(poll_sleep)
        return results
