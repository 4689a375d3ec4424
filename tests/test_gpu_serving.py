"""Serving on the GPU: a small request stream (uneven sizes, parameter sweep,
Poisson arrivals) through the real batch with continuous admission, a tuned
per-size tiling, and the check that every served result equals the same run
alone, bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from mysph_amd import config as C
from mysph_amd import serving
from mysph_amd.sphgpu import FINISHED, Batch

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "rho")


def test_served_results_equal_runs_alone():
    reqs = serving.sweep_requests(10, lo=2000, hi=30000, steps=120)
    serving.poisson_arrivals(reqs, rate=40.0)
    tuner = serving.Tuner(bounds=(8000,), table=[(3, 0.15), (2, 0.25)])
    total = sum(r.case["x"].shape[0] for r in reqs)
    print(f"serving 10 runs, {total} particles", flush=True)
    # room for about a third of the stream at once: admission has to wait
    with Batch(2, C.MIXED, slots=total // 3 + 40000, cells=4_000_000, max_runs=16) as b:
        res = serving.Server(b, chunk=40, tuner=tuner, keep=FIELDS).serve(reqs)
    assert len(res) == 10 and all(r.status == FINISHED for r in res)
    st = serving.latency_stats(res)
    assert st["runs"] == 10 and st["p99_s"] >= st["p50_s"] > 0
    by_id = {r.rid: r for r in res}
    for req in reqs:
        win, skin = tuner.choice(req.n_fluid)
        with Batch(2, C.MIXED, slots=req.case["x"].shape[0] + 256, cells=1_000_000, max_runs=1) as b:
            h = b.add(req.case, req.steps, win=win, skin_frac=skin)
            b.advance(req.steps)
            assert b.status(h)[1] == FINISHED
            for k in FIELDS:
                assert np.array_equal(by_id[req.rid].fields[k], b.get(h, k)), (req.rid, k)


def test_tuner_picks_a_measured_candidate():
    tuner = serving.Tuner(bounds=(10000,))

    def rep(k):
        return serving.dam_break_request_case(6000 if k == 0 else 40000, 1.2, 0.1)

    def factory(slots):
        return Batch(2, C.MIXED, slots=slots, cells=8_000_000, max_runs=64)

    report = tuner.tune(factory, rep, candidates=((3, 0.15), (2, 0.25)), fill=200_000, warmup=3, steps=6)
    for k in (0, 1):
        assert len(report[k]) == 2 and all(t[2] > 0 for t in report[k])
        assert tuner.table[k] == report[k][0][:2]
