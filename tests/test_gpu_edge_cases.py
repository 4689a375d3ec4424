"""Edge cases of the hot path against the oracle, on both pipelines (the
fused fast path and the general path) and in both precisions where they
apply:

  - particles with State == -1, which BGGS skips as targets and sources
    (asymmetricSearch.f90:89-154, :223-226);
  - a particle that leaves the fixed search grid (its key is clamped to the
    border and its pairs take the FP64 test), while the reference rebuilds
    its grid from the bounding box every step;
  - an isolated particle with no neighbour, and a run with a single real
    particle among the walls;
  - a context with spare capacity, uploaded twice with different sets;
  - a particle set that outgrows the neighbour lists planned at upload: the
    fast path walks the cell window of such a particle and stays exact; the
    general path fails with SPHGPU_E_PAIR_CAPACITY (the reference grows its
    list, asymmetricSearch.f90:239-245), and the state uploads again.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from mysph_amd import cases
from mysph_amd import config as C
from mysph_amd.sphgpu import Simulation, SphGpuError
from tests.parity import assert_close

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "rho", "p", "e")


def _compare(case, cfg, steps, tol=1e-10, general=False, monkeypatch=None):
    nt = case["n_real"]
    if general:
        monkeypatch.setenv("SPHGPU_FORCE_GENERAL", "1")
    orc = oracle.OracleSim(case, cfg)
    orc.run(steps)
    with Simulation(case, cfg) as sim:
        sim.advance(steps)
        for k in FIELDS:
            assert_close(k, sim.get(k)[:nt], orc[k][:nt], tol)
        return sim.pairs(), orc


@pytest.mark.parametrize("general", [False, True])
def test_state_minus_one_particles_are_skipped(general, monkeypatch):
    case = cases.dam_break_2d(dx=0.04)
    rng = np.random.default_rng(11)
    pick = rng.choice(case["n_real"], 40, replace=False)
    case["state"][pick] = -1
    cfg = case["config"].copy()
    print(f"State -1 on 40 particles, general={general}", flush=True)
    (off, nbr), _ = _compare(case, cfg, 30, general=general, monkeypatch=monkeypatch)
    cnt = np.diff(off)
    assert np.all(cnt[pick] == 0)                      # not targets
    assert not np.isin(nbr, pick).any()                # not sources


@pytest.mark.parametrize("precision", [C.FP64, C.MIXED])
def test_particle_leaving_the_search_grid(precision):
    case = cases.dam_break_2d(dx=0.04)
    nt = case["n_real"]
    # the top fluid particle is thrown straight up at 80 m/s: it leaves the
    # tank over its walls and the grid planned at upload (walls to 3 m plus a
    # margin) within a few hundred steps
    i = int(np.argmax(case["x"][:nt, 1]))
    case["v"][i] = (0.0, 80.0)
    cfg = case["config"].copy()
    cfg.precision = precision
    steps = 300
    print(f"escaping particle, precision={precision}", flush=True)
    orc = oracle.OracleSim(case, cfg)
    orc.run(steps)
    with Simulation(case, cfg) as sim:
        gx = sim.grid_info()
        sim.advance(steps)
        x = sim.get("x")
        assert x[i, 1] > case["x"][:, 1].max() + 1.0    # far outside the initial box
        tol = 1e-10 if precision == C.FP64 else None
        if tol:
            for k in FIELDS:
                assert_close(k, sim.get(k)[:nt], orc[k][:nt], tol)
        else:
            # mixed: the same neighbour sets, and the state close to FP64's
            assert_close("x", x[:nt], orc["x"][:nt], 1e-6)
    assert gx is not None


def test_isolated_and_single_particles(monkeypatch):
    case = cases.dam_break_2d(dx=0.04)
    nt = case["n_real"]
    # one fluid particle far above the column: no neighbour at all
    case["x"][nt - 1] = (0.5, 2.9)
    (off, nbr), _ = _compare(case, case["config"].copy(), 20)
    assert np.diff(off)[nt - 1] == 0
    # a single real particle among the walls
    one = dict(case)
    keep = np.r_[[0], np.arange(nt, case["x"].shape[0])]
    for k in ("x", "v", "mass", "rho", "p", "e", "c", "h", "type", "state", "boundary"):
        one[k] = np.ascontiguousarray(case[k][keep])
    one["n_real"] = 1
    _compare(one, one["config"].copy(), 20)


def test_spare_capacity_and_second_upload():
    """A context keeps its config; a second upload replaces the particle set
    (here a larger one) and restarts at step 1."""
    a = cases.dam_break(2, 0.04, column=(0.8, 1.6))
    b = cases.dam_break_2d(dx=0.04)
    cfg = a["config"].copy()
    cap = b["x"].shape[0] + 1000
    ref = {}
    for name, case in (("a", a), ("b", b)):
        orc = oracle.OracleSim(case, cfg)
        orc.run(25)
        ref[name] = orc
    with Simulation(a, cfg, capacity=cap) as sim:
        sim.advance(25)
        nt = a["n_real"]
        for k in FIELDS:
            assert_close("first " + k, sim.get(k)[:nt], ref["a"][k][:nt], 1e-10)
        sim.n_real, sim.n_dummy = b["n_real"], b["n_dummy"]
        sim.n = b["x"].shape[0]
        sim.upload(b)
        assert sim.step == 1
        sim.advance(25)
        nt = b["n_real"]
        for k in FIELDS:
            assert_close("second " + k, sim.get(k)[:nt], ref["b"][k][:nt], 1e-10)


def _converging_cloud():
    """A 20 x 20 cloud at twice the lattice spacing (4 neighbours each) that
    converges ballistically: no EOS (pressure stays at its uploaded 0), no
    viscosity, no gravity, v = -20 (x - c). Near half its spacing every
    particle has about 20 neighbours, more than the 12 the lists are planned
    for at upload (1.5x the largest count + 8)."""
    dx, m = 0.02, 20
    g = np.arange(m) * 2.0 * dx
    X, Y = np.meshgrid(g, g, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel()], axis=1) + 1.0
    n = x.shape[0]
    v = -20.0 * (x - x.mean(axis=0))
    cfg = C.dam_break_config(2, 1.0, 1.2 * dx)
    cfg.gravity = 0
    cfg.artificial_viscosity = 0
    cfg.mat[0].eos = C.EOS_NONE
    return dict(x=x, v=v, mass=np.full(n, 1000.0 * dx * dx), rho=np.full(n, 1000.0),
                p=np.zeros(n), e=np.zeros(n), c=np.zeros(n), h=np.full(n, 1.2 * dx),
                type=np.full(n, 2, dtype=np.int32), state=np.zeros(n, dtype=np.int32),
                boundary=np.zeros(n, dtype=np.int32), n_real=n, n_dummy=0, dim=2, config=cfg)


def test_outgrown_candidate_lists_stay_exact():
    """Fast path: a particle whose candidates outgrow its list walks its cell
    window instead, so the steps stay exact (the reference grows its list,
    asymmetricSearch.f90:239-245)."""
    case = _converging_cloud()
    nt = case["n_real"]
    print("outgrown candidate lists, fast path", flush=True)
    orc = oracle.OracleSim(case)
    orc.run(300)
    with Simulation(case) as sim:
        sim.advance(300)
        assert sim.neighbour_stats()["max"] > 12     # past the planned list
        for k in FIELDS:
            assert_close(k, sim.get(k)[:nt], orc[k][:nt], 1e-10)


def test_outgrown_lists_fail_loudly(monkeypatch):
    """General path: the exact list planned at upload is outgrown, and the
    step fails with SPHGPU_E_PAIR_CAPACITY instead of dropping pairs.
    Uploading the state again plans larger lists."""
    monkeypatch.setenv("SPHGPU_FORCE_GENERAL", "1")
    case = _converging_cloud()
    cfg = case["config"]
    print("outgrown lists, general path", flush=True)
    err = None
    with Simulation(case, cfg) as sim:
        for _ in range(20):
            try:
                sim.advance(20)
            except SphGpuError as e:
                err = e
                break
        assert err is not None, "the compressed cloud never outgrew its lists"
        assert err.code == -3 and "upload" in str(err), str(err)
        state = {k: sim.get(k) for k in ("x", "v", "rho")}
    again = dict(case, **state)
    with Simulation(again, cfg) as sim:
        sim.advance(10)
        assert np.all(np.isfinite(sim.get("x")))
