"""The serving scheduler on the CPU, against a stand-in batch with the
admission and step semantics of sphgpu_batch_* and a simulated clock: FIFO
admission by room, no over-commitment, every request finishes, failures are
reported, and the latency statistics add up. Also the request stream."""
from __future__ import annotations

import math

import numpy as np
import pytest

from mysph_amd import serving


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(0.0, dt)


class FakeBatch:
    """Room = slot count; a step costs `sec_per_ps` per resident particle."""

    def __init__(self, slots, clock, sec_per_ps=1e-9, fail=()):
        self.slots, self.clock, self.sec = slots, clock, sec_per_ps
        self.runs = {}
        self.next = 0
        self.peak = 0
        self.fail = set(fail)
        self.order = []

    def add(self, case, nsteps, win=0, skin_frac=0.0):
        n = case["n"]
        if sum(r["n"] for r in self.runs.values()) + n > self.slots:
            return None
        h = self.next
        self.next += 1
        self.runs[h] = dict(n=n, step=1, target=nsteps, state=0, rid=case["rid"])
        self.order.append(case["rid"])
        self.peak = max(self.peak, sum(r["n"] for r in self.runs.values()))
        return h

    def advance(self, k):
        left = k
        while left > 0:
            live = [r for r in self.runs.values() if r["state"] == 0]
            if not live:
                break
            m = min([left] + [r["target"] - r["step"] + 1 for r in live])
            self.clock.t += m * self.sec * sum(r["n"] for r in live)
            for r in live:
                r["step"] += m
                if r["rid"] in self.fail:
                    r["state"] = -2
                elif r["step"] > r["target"]:
                    r["state"] = 1
            left -= m

    def status(self, h):
        r = self.runs[h]
        return r["step"], r["state"]

    def get(self, h, name):
        return np.zeros(1)

    def remove(self, h):
        del self.runs[h]


def fake_requests(sizes, steps=100, arrivals=None):
    reqs = []
    for k, n in enumerate(sizes):
        r = serving.Request(k, n, 1.0, 0.1, steps)
        r.case = {"n": n, "rid": k, "n_real": n}
        r.arrival = 0.0 if arrivals is None else arrivals[k]
        reqs.append(r)
    return reqs


def test_every_request_finishes_within_room():
    clock = FakeClock()
    b = FakeBatch(slots=1000, clock=clock)
    reqs = fake_requests([300, 500, 200, 900, 100, 400, 600], arrivals=[0, 0.1, 0.1, 0.2, 5, 5, 9])
    res = serving.Server(b, chunk=10).serve(reqs, clock=clock, sleep=clock.sleep)
    assert sorted(r.rid for r in res) == list(range(7))
    assert b.peak <= 1000
    assert all(r.status == 1 for r in res)
    assert all(r.admitted >= r.arrival - 1e-12 and r.done > r.admitted for r in res)
    # FIFO: admission order is arrival order (ties in request order)
    assert b.order == [0, 1, 2, 3, 4, 5, 6]


def test_head_of_line_waits_for_room():
    clock = FakeClock()
    b = FakeBatch(slots=1000, clock=clock)
    reqs = fake_requests([800, 700, 100], steps=50)
    res = {r.rid: r for r in serving.Server(b, chunk=10).serve(reqs, clock=clock, sleep=clock.sleep)}
    # 700 does not fit beside 800; 100 (behind it) waits too
    assert res[1].admitted >= res[0].done - 1e-12
    assert res[2].admitted >= res[1].admitted


def test_failed_run_is_reported_and_others_finish():
    clock = FakeClock()
    b = FakeBatch(slots=10_000, clock=clock, fail={2})
    res = {r.rid: r for r in serving.Server(b, chunk=25).serve(fake_requests([100] * 5), clock=clock,
                                                               sleep=clock.sleep)}
    assert res[2].status == -2
    assert all(res[k].status == 1 for k in (0, 1, 3, 4))
    st = serving.latency_stats(list(res.values()))
    assert st["runs"] == 4 and st["failed"] == 1


def test_request_larger_than_arena_raises():
    clock = FakeClock()
    b = FakeBatch(slots=100, clock=clock)
    with pytest.raises(RuntimeError):
        serving.Server(b).serve(fake_requests([500]), clock=clock, sleep=clock.sleep)


def test_latency_stats_arithmetic():
    res = [serving.Result(k, 1000, 10, float(k), float(k), float(k) + 1.0 + k, 1) for k in range(100)]
    st = serving.latency_stats(res)
    lat = np.array([1.0 + k for k in range(100)])
    assert st["p50_s"] == pytest.approx(np.percentile(lat, 50))
    assert st["p99_s"] == pytest.approx(np.percentile(lat, 99))
    assert st["particle_steps"] == 100 * 1000 * 10
    assert st["span_s"] == pytest.approx(99.0 + 1.0 + 99.0)
    assert st["throughput"] == pytest.approx(st["particle_steps"] / st["span_s"])


def test_poisson_arrivals_rate():
    reqs = fake_requests([10] * 4000)
    serving.poisson_arrivals(reqs, rate=50.0)
    t = np.array([r.arrival for r in reqs])
    assert t[0] == 0.0 and np.all(np.diff(t) >= 0)
    assert 3999 / t[-1] == pytest.approx(50.0, rel=0.05)
    serving.poisson_arrivals(reqs, rate=math.inf)
    assert all(r.arrival == 0.0 for r in reqs)


def test_sweep_requests_distribution_and_sizes():
    reqs = serving.sweep_requests(256, build=False)
    n = np.array([r.n_target for r in reqs])
    assert n.min() >= 5000 and n.max() <= 500000
    # log-uniform: about half below the geometric mean
    assert 0.4 < np.mean(n < math.sqrt(5000 * 500000)) < 0.6
    assert all(0.5 <= r.aspect <= 2.0 and 0.05 <= r.alpha <= 0.2 for r in reqs)
    again = serving.sweep_requests(256, build=False)
    assert [r.n_target for r in again] == list(n)
    # the built case has about the requested fluid count
    for r in reqs[:6]:
        c = serving.dam_break_request_case(r.n_target, r.aspect, r.alpha)
        assert abs(c["n_real"] - r.n_target) < 0.03 * r.n_target + 100
        assert c["config"].av_alpha == r.alpha


def test_tuner_classes():
    t = serving.Tuner(bounds=(100, 1000))
    assert [t.size_class(n) for n in (50, 100, 101, 1000, 5000)] == [0, 0, 1, 1, 2]
    t.table = [(3, 0.15), (2, 0.15), (2, 0.3)]
    assert t.choice(5000) == (2, 0.3)
