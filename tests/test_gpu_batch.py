"""Batches of runs (sphgpu_batch_*): one arena, one launch per pipeline stage
for all resident runs.

Bars:
  * a run alone in a batch equals the same run in a single-run context, bit
    for bit;
  * a run in a mix (different sizes, parameters, step counts, admitted at
    different times) equals the same run alone, bit for bit;
  * a run that fails (non-finite state) stops alone and the others still
    equal their runs alone;
  * admission reports "no room" without changing the batch, and freed room
    is reused.
"""
from __future__ import annotations

import numpy as np
import pytest

from mysph_amd import cases
from mysph_amd import config as C
from mysph_amd.sphgpu import FINISHED, RUNNING, Batch, Simulation

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "rho", "p", "e")


def run_case(dx, column=(1.0, 2.0), tank=(4.0, 3.0), alpha=0.1, seed_v=None, precision=C.MIXED):
    c = cases.dam_break(2, dx, column=column, tank=tank, alpha=alpha, precision=precision)
    if seed_v is not None:
        nt = c["n_real"]
        rng = np.random.default_rng(seed_v)
        c["v"] = np.zeros_like(c["x"])
        c["v"][:nt] = rng.normal(0.0, 1.0, (nt, 2))
    return c


def alone(case, nsteps, precision):
    with Batch(2, precision, slots=case["x"].shape[0] + 512, cells=400000, max_runs=4) as b:
        r = b.add(case, nsteps)
        b.advance(nsteps)
        step, state = b.status(r)
        assert state == FINISHED and step == nsteps + 1
        return {k: b.get(r, k) for k in FIELDS}


def assert_same(got, ref, label):
    for k in FIELDS:
        assert np.array_equal(got[k], ref[k]), f"{label}: {k} differs"


@pytest.mark.parametrize("precision", [C.FP64, C.MIXED])
def test_batch_alone_equals_single_context(precision):
    case = run_case(0.04, seed_v=3, precision=precision)
    print(f"batch vs context, precision {precision}", flush=True)
    ref = alone(case, 60, precision)
    cfg = case["config"].copy()
    cfg.precision = precision
    with Simulation(case, cfg) as sim:
        sim.advance(60)
        assert_same({k: sim.get(k) for k in FIELDS}, ref, "single context")


@pytest.mark.parametrize("precision", [C.FP64, C.MIXED])
def test_run_in_mix_equals_run_alone(precision):
    specs = [
        ("A", run_case(0.05, precision=precision), 40),
        ("B", run_case(0.04, column=(0.8, 1.2), tank=(3.0, 2.0), alpha=0.15, precision=precision), 60),
        ("C", run_case(0.06, column=(0.5, 1.0), tank=(2.0, 1.5), seed_v=8, precision=precision), 30),
        ("D", run_case(0.045, column=(1.0, 1.0), tank=(3.0, 2.0), alpha=0.06, precision=precision), 45),
    ]
    print(f"mix of 4 runs, precision {precision}", flush=True)
    refs = {name: alone(case, n, precision) for name, case, n in specs}
    total = sum(c["x"].shape[0] for _, c, _ in specs)
    with Batch(2, precision, slots=total + 4096, cells=2000000, max_runs=8) as b:
        ids = {}
        for name, case, n in specs[:3]:
            ids[name] = b.add(case, n)
        b.advance(20)
        ids["D"] = b.add(specs[3][1], specs[3][2])       # admitted while the others run
        for _ in range(4):
            b.advance(17)
        for name, case, n in specs:
            step, state = b.status(ids[name])
            assert state == FINISHED and step == n + 1, (name, step, state)
            assert_same({k: b.get(ids[name], k) for k in FIELDS}, refs[name], f"run {name}")


def test_failed_run_stops_alone():
    good = run_case(0.05)
    other = run_case(0.04, column=(0.8, 1.2), tank=(3.0, 2.0))
    bad = run_case(0.06)
    bad["v"] = np.zeros_like(bad["x"])
    bad["v"][0, 0] = np.nan                       # non-finite state after the first step
    ref_good = alone(good, 30, C.MIXED)
    ref_other = alone(other, 30, C.MIXED)
    print("mix with a failing run", flush=True)
    with Batch(2, C.MIXED, slots=60000, cells=2000000, max_runs=4) as b:
        g = b.add(good, 30)
        x = b.add(bad, 30)
        o = b.add(other, 30)
        b.advance(30)
        step, state = b.status(x)
        assert state == -2, (step, state)             # SPHGPU_E_NONFINITE
        assert "non-finite" in b.run_error(x)
        for r, ref, label in ((g, ref_good, "good"), (o, ref_other, "other")):
            assert b.status(r) == (31, FINISHED)
            assert_same({k: b.get(r, k) for k in FIELDS}, ref, label)


def test_admission_by_arena_room():
    case = run_case(0.1)
    n = case["x"].shape[0]
    nslot = (n + 255) // 256 * 256
    with Batch(2, C.MIXED, slots=2 * nslot, cells=200000, max_runs=8) as b:
        r1 = b.add(case, 10)
        r2 = b.add(case, 20)
        assert r1 is not None and r2 is not None
        before = b.info()
        assert b.add(case, 10) is None                 # no slot range left
        assert b.info() == before                      # nothing changed
        b.advance(10)
        assert b.status(r1) == (11, FINISHED)
        assert b.status(r2) == (11, RUNNING)
        b.remove(r1)
        r3 = b.add(case, 10)                           # reuses the freed range
        assert r3 is not None
        b.advance(10)
        assert b.status(r2) == (21, FINISHED) and b.status(r3) == (11, FINISHED)
        ms, steps, ps = b.timing()
        assert steps == 10 and ps == 2 * 10 * case["n_real"] and ms > 0
        # results of the reused range equal a fresh run
        assert_same({k: b.get(r3, k) for k in FIELDS}, alone(case, 10, C.MIXED), "reused")
