"""Serving many independent SPH runs on one GPU with continuous batching.

Requests are dam-break runs of uneven size (a parameter sweep). Each request
is admitted into a resident batch (sphgpu_batch_*) as soon as the arena has
room for it, advances together with every other resident run (one launch per
pipeline stage covers all of them), and leaves the batch the chunk in which
it reaches its last step, freeing room for the next request. Results of a
run do not depend on which other runs shared the batch (the batch guarantees
it bit for bit), so serving is a pure scheduling question.

The reference runs one case per process (src/main.f90:16-22); this module has
no counterpart there.

Pieces:
  ``sweep_requests``   the request stream: counts log-uniform in [lo, hi],
                       column aspect U[0.5, 2], alpha U[0.05, 0.2], seeded
  ``Tuner``            per-size choice of the neighbour-search tiling (cells
                       per cutoff, candidate skin), measured on the GPU
  ``Server``           FIFO admission by arena room, chunked advance,
                       completion bookkeeping
  ``latency_stats``    p50 / p99 time to completion and aggregate throughput
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import cases
from . import config as C

SEED_SWEEP = 20261204


@dataclass
class Request:
    rid: int
    n_target: int           # requested fluid particle count
    aspect: float           # column height / width
    alpha: float            # Monaghan alpha
    steps: int
    arrival: float = 0.0    # seconds after the start of the stream
    case: dict | None = None

    @property
    def n_fluid(self) -> int:
        return int(self.case["n_real"]) if self.case is not None else self.n_target


@dataclass
class Result:
    rid: int
    n_fluid: int
    steps: int
    arrival: float
    admitted: float
    done: float
    status: int                      # 1 finished, < 0 the run's failure status
    fields: dict = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def dam_break_request_case(n: int, aspect: float, alpha: float, precision: int = C.MIXED) -> dict:
    """2D dam break with about n fluid particles: a 1 m wide column of height
    `aspect` m in a tank 4 column widths long and 1.5 column heights deep;
    dx = sqrt(area / n)."""
    W, H = 1.0, float(aspect)
    dx = math.sqrt(W * H / n)
    return cases.dam_break(2, dx, column=(W, H), tank=(4.0 * W, 1.5 * H), alpha=alpha,
                           precision=precision)


def sweep_requests(nruns: int = 256, lo: int = 5000, hi: int = 500000, steps: int = 2000,
                   seed: int = SEED_SWEEP, precision: int = C.MIXED, build: bool = True) -> list:
    """The serving workload: counts log-uniform in [lo, hi], aspect U[0.5, 2],
    alpha U[0.05, 0.2], `steps` steps each."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(math.log(lo), math.log(hi), nruns)).astype(int)
    aspects = rng.uniform(0.5, 2.0, nruns)
    alphas = rng.uniform(0.05, 0.2, nruns)
    reqs = []
    for k in range(nruns):
        r = Request(k, int(counts[k]), float(aspects[k]), float(alphas[k]), int(steps))
        if build:
            r.case = dam_break_request_case(r.n_target, r.aspect, r.alpha, precision)
        reqs.append(r)
    return reqs


def poisson_arrivals(reqs: list, rate: float, seed: int = SEED_SWEEP + 1) -> list:
    """Arrival times of a Poisson process of `rate` requests/s (rate = inf:
    all at time 0), in request order."""
    if not math.isfinite(rate):
        for r in reqs:
            r.arrival = 0.0
        return reqs
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
    t -= t[0]
    for r, a in zip(reqs, t):
        r.arrival = float(a)
    return reqs


def latency_stats(results: list) -> dict:
    """p50 / p99 time to completion (arrival to done) and the aggregate
    throughput of the finished runs over the span of the stream."""
    ok = [r for r in results if r.status == 1]
    if not ok:
        return {"runs": 0}
    lat = np.array([r.latency for r in ok])
    span = max(r.done for r in ok) - min(r.arrival for r in ok)
    work = sum(r.n_fluid * r.steps for r in ok)
    return {
        "runs": len(ok),
        "failed": len(results) - len(ok),
        "p50_s": float(np.percentile(lat, 50)),
        "p99_s": float(np.percentile(lat, 99)),
        "mean_s": float(lat.mean()),
        "max_s": float(lat.max()),
        "span_s": float(span),
        "particle_steps": int(work),
        "throughput": float(work / span) if span > 0 else 0.0,
        "queue_wait_p50_s": float(np.percentile([r.admitted - r.arrival for r in ok], 50)),
    }


class Tuner:
    """Per-size choice of (cells per cutoff, skin / cutoff) for the batched
    neighbour search. Both leave the interacting pairs unchanged; the choice
    is a pure function of the run's size class once tuned, so a run takes
    the same path alone and in a mix.

    ``tune`` times each candidate on a batch of replicas of a representative
    run (about `fill` particles in all) and keeps the fastest per class. The
    candidates span thin skins (fewer candidates per particle, more
    rebuilds) to thick ones and coarse to fine cells; when one of them wins
    every class, the report says so with the measured margins."""

    DEFAULT_CANDIDATES = ((3, 0.15), (3, 0.08), (4, 0.1), (2, 0.15), (3, 0.25))

    def __init__(self, bounds=(20000, 80000, 250000), table=None):
        self.bounds = tuple(bounds)          # upper fluid counts of the classes; last is open
        self.table = list(table) if table else [(0, 0.0)] * (len(self.bounds) + 1)

    def size_class(self, n_fluid: int) -> int:
        for k, b in enumerate(self.bounds):
            if n_fluid <= b:
                return k
        return len(self.bounds)

    def choice(self, n_fluid: int):
        return self.table[self.size_class(n_fluid)]

    def tune(self, batch_factory, representative, candidates=DEFAULT_CANDIDATES, fill=2_000_000,
             warmup=5, steps=20):
        """representative(k) -> case of class k. Returns {class: [(win, skin, ns per
        particle-step), ...]}."""
        report = {}
        for k in range(len(self.bounds) + 1):
            case = representative(k)
            n = case["x"].shape[0]
            reps = max(1, fill // n)
            timings = []
            for win, skin in candidates:
                with batch_factory(reps * ((n + 255) // 256 * 256) + 256) as b:
                    handles = [b.add(case, warmup + steps, win=win, skin_frac=skin) for _ in range(reps)]
                    if any(h is None for h in handles):
                        continue
                    b.advance(warmup)
                    b.advance(steps)
                    ms, st, ps = b.timing()
                    timings.append((win, skin, 1e6 * ms / max(ps, 1)))
            timings.sort(key=lambda t: t[2])
            report[k] = timings
            if timings:
                self.table[k] = (timings[0][0], timings[0][1])
        return report


class Server:
    """Continuous batching over one resident batch.

    Each loop: move arrived requests to the queue, admit from the head of the
    queue while the arena has room (FIFO: a request that does not fit waits,
    and so do those behind it), advance every resident run by up to `chunk`
    steps, and retire the runs that finished (or failed) in that chunk."""

    def __init__(self, batch, chunk: int = 50, tuner: Tuner | None = None, keep=(), keep_ids=None):
        self.batch = batch
        self.chunk = int(chunk)
        self.tuner = tuner
        self.keep = tuple(keep)          # fields to download from each finished run
        self.keep_ids = None if keep_ids is None else set(keep_ids)   # only these requests
        self.advance_s = 0.0
        self.chunks = 0

    def _admit(self, req: Request):
        win, skin = (0, 0.0)
        if self.tuner is not None:
            win, skin = self.tuner.choice(req.n_fluid)
        return self.batch.add(req.case, req.steps, win=win, skin_frac=skin)

    def serve(self, requests: list, clock=time.perf_counter, sleep=time.sleep) -> list:
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        queue = deque()
        active = {}
        results = []
        t0 = clock()
        while pending or queue or active:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            while queue:
                h = self._admit(queue[0])
                if h is None:
                    break
                req = queue.popleft()
                active[h] = (req, clock() - t0)
            if not active:
                if queue:
                    raise RuntimeError(f"request {queue[0].rid} does not fit an empty batch")
                if pending:
                    sleep(max(0.0, pending[0].arrival - (clock() - t0)))
                continue
            a = clock()
            self.batch.advance(self.chunk)
            self.advance_s += clock() - a
            self.chunks += 1
            now = clock() - t0
            for h in list(active):
                step, state = self.batch.status(h)
                if state == 0:
                    continue
                req, admitted = active.pop(h)
                fields = ({k: self.batch.get(h, k) for k in self.keep}
                          if self.keep_ids is None or req.rid in self.keep_ids else {})
                self.batch.remove(h)
                results.append(Result(req.rid, req.n_fluid, req.steps, req.arrival, admitted, now,
                                      state, fields))
        return results
