"""Serving on the GPU: a small request stream (uneven sizes, parameter sweep,
Poisson arrivals) through the real batch with continuous admission, a tuned
per-size tiling, and the check that every served result equals the same run
alone, bit for bit; batched runs against the oracle directly, in both
precisions, also when their candidate lists overflow."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from mysph_amd import config as C
from mysph_amd import serving
from mysph_amd.sphgpu import FINISHED, Batch
from tests.parity import assert_close, rel_err

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


def _batch_cases(precision):
    return [serving.dam_break_request_case(n, aspect, alpha, precision=precision)
            for n, aspect, alpha in ((1500, 1.0, 0.1), (4000, 1.5, 0.15), (2500, 0.7, 0.08))]


def _check_vs_oracle(b, h, case, steps, precision):
    nt = case["n_real"]
    orc = oracle.OracleSim(case)
    orc.run(steps)
    if precision == C.FP64:
        for k in ("x", "v", "rho", "p", "e"):
            assert_close(k, b.get(h, k)[:nt], orc[k][:nt], 1e-9)
    else:
        rho = b.get(h, "rho")[:nt]
        e_rho = float(np.max(np.abs(rho - orc["rho"][:nt]) / orc["rho"][:nt]))
        e_x = rel_err(b.get(h, "x")[:nt], orc["x"][:nt])
        print(f"mixed batched run after {steps} steps: rho {e_rho:.2e}, x {e_x:.2e}", flush=True)
        assert e_rho <= 1e-4


@pytest.mark.parametrize("precision", [C.FP64, C.MIXED])
def test_batched_run_vs_oracle(precision):
    """A run served beside two others against the oracle directly (not only
    through the single-run context): FP64 within 1e-9 after 100 steps (the
    single run's bar), mixed-precision density within 1e-4."""
    cs = _batch_cases(precision)
    total = sum(c["x"].shape[0] for c in cs)
    print(f"batch of 3 runs, precision {precision}: run 1 against the oracle", flush=True)
    with Batch(2, precision, slots=total + 3 * 256, cells=4_000_000, max_runs=4) as b:
        hs = [b.add(c, 100) for c in cs]
        b.advance(100)
        assert all(b.status(h)[1] == FINISHED for h in hs)
        _check_vs_oracle(b, hs[1], cs[1], 100, precision)


@pytest.mark.parametrize("precision", [C.FP64, C.MIXED])
def test_batched_candidate_overflow_continues(precision):
    """A batch whose candidate lists hold 8 rows (a 2D particle has about 20
    candidates): the particles walk their cell windows, their runs rebuild
    every step, and every run finishes and matches the oracle after 60
    steps (FP64 1e-9, mixed-precision density 1e-4)."""
    cs = _batch_cases(precision)[:2]
    total = sum(c["x"].shape[0] for c in cs)
    print(f"batch with 8 candidate rows, precision {precision}", flush=True)
    with Batch(2, precision, slots=total + 2 * 256, cells=4_000_000, max_runs=4, verlet_cap=8) as b:
        hs = [b.add(c, 60) for c in cs]
        b.advance(60)
        assert all(b.status(h)[1] == FINISHED for h in hs), [b.status(h) for h in hs]
        for h, c in zip(hs, cs):
            _check_vs_oracle(b, h, c, 60, precision)
