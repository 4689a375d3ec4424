"""General path (solids, kernel correction, physical viscosity, summation
density, evolving h, SPH 1/3) on candidate lists: the cell sort and the
candidate list (every j within cutoff + skin) are rebuilt only when some
particle has moved skin/2, each step filters the candidates into the exact
list with the reference's rule (asymmetricSearch.f90:209-237), and groups of
steps replay a captured graph.

  - against the same runs on exact lists from a cell sweep every step
    (SPHGPU_NO_VERLET=1): the pair sets are the same, only the pair order
    differs, so the fields agree to rounding (1e-12; 1e-10 for the 3D beam,
    whose small pressure differs at 1.6e-12 after 40 steps; the
    artificial-stress beam's stress rotation amplifies last bits, so its
    bound is 10x the oracle's own response to a 1e-15 perturbation);
  - rebuild-heavy runs (random velocities of ~1.5 m/s) against the oracle,
    with and without an evolving h;
  - particles whose candidates outgrow their list (SPHGPU_VERLET_CAP) walk
    their cell window and stay exact;
  - split calls and graph replays equal one eager call bit for bit;
  - a run past the small-run size (one lane per fluid target, solid
    targets on a second stream, the mixed filter inside the strain pass)
    against the oracle, and its fused filter against the separate one.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from mysph_amd import cases
from mysph_amd import config as C
from mysph_amd.sphgpu import Simulation
from tests.parity import assert_close, rel_err

pytestmark = pytest.mark.gpu

STATE_FIELDS = ("x", "v", "rho", "p", "e")


def _beam(which, asw):
    from tests.golden.make_golden import beam_velocity
    case = cases.reference_can_beam() if which == "2d" else cases.reference_can_beam_3d()
    case["v"] = beam_velocity(case["x"])
    case["config"] = case["config"].copy()
    case["config"].artificial_stress = asw
    return case


def _ref_dam_break(**switches):
    case = cases.reference_dam_break()
    case["c"][case["n_real"]:] = 50.0
    cfg = case["config"].copy()
    for k, v in switches.items():
        setattr(cfg, k, v)
    case["config"] = cfg
    return case


def _stirred(viscosity=1, sle=0, seed=11):
    """2D dam break (dx = 0.04) with random velocities: particles cross the
    skin many times in 60 steps, so the candidate list is rebuilt often."""
    case = cases.dam_break_2d(dx=0.04)
    nt = case["n_real"]
    rng = np.random.default_rng(seed)
    case["v"] = np.zeros_like(case["x"])
    case["v"][:nt] = rng.normal(0.0, 1.5, (nt, 2))
    cfg = case["config"].copy()
    cfg.viscosity = viscosity
    cfg.sle = sle
    case["config"] = cfg
    return case


CASES = {
    "beam2d": (lambda: _beam("2d", 0), 1e-12, ("stress",)),
    "beam2d_as": (lambda: _beam("2d", 1), None, ("stress",)),
    "beam3d": (lambda: _beam("3d", 0), 1e-10, ("stress",)),
    "gate": (lambda: cases.dam_break_gate_3d(dx_f=0.005, dx_s=0.0025, width=0.03), 1e-12, ("stress",)),
    "sle2": (lambda: _ref_dam_break(sle=2), 1e-12, ("h",)),
    "sph3": (lambda: _ref_dam_break(pa_sph=3), 1e-12, ()),
    "shepard": (lambda: _ref_dam_break(summation_density=1, normalize_density=1), 1e-12, ()),
}


@pytest.mark.parametrize("name", list(CASES))
def test_candidate_lists_equal_cell_sweep_lists(name, monkeypatch):
    make, tol, extra = CASES[name]
    case = make()
    nt = case["n_real"]
    print("general path, candidate vs exact lists:", name, flush=True)
    tols = {k: tol for k in STATE_FIELDS + extra}
    if tol is None:
        orc = oracle.OracleSim(case)
        orc.run(40)
        pert = dict(case)
        pert["v"] = case["v"] * (1 + 1e-15 * np.random.default_rng(1).standard_normal(case["v"].shape))
        o2 = oracle.OracleSim(pert)
        o2.run(40)
        tols = {k: max(1e-10, 10 * rel_err(o2[k][:nt], orc[k][:nt])) for k in tols}
    with Simulation(case) as a:
        a.advance(40)
        ra = {k: a.get(k)[:nt] for k in STATE_FIELDS + extra}
        pa = a.neighbour_stats()["pairs"]
    monkeypatch.setenv("SPHGPU_NO_VERLET", "1")
    with Simulation(case) as b:
        b.advance(40)
        assert b.neighbour_stats()["pairs"] == pa
        for k in STATE_FIELDS + extra:
            assert_close(f"{name} {k}", b.get(k)[:nt], ra[k], tols[k])


@pytest.mark.parametrize("sle", [0, 2], ids=["fixed_h", "sle2"])
def test_rebuilds_keep_the_general_path_exact(sle):
    case = _stirred(sle=sle)
    nt = case["n_real"]
    print("general path, stirred dam break, sle", sle, flush=True)
    orc = oracle.OracleSim(case)
    orc.run(60)
    with Simulation(case) as sim:
        sim.advance(60)
        for k in STATE_FIELDS + (("h",) if sle else ()):
            assert_close(k, sim.get(k)[:nt], orc[k][:nt], 1e-8)


@pytest.mark.parametrize("general", [False, True], ids=["fast_path", "general_path"])
def test_outgrown_candidate_rows_walk_the_cell_window(general, monkeypatch):
    """With 8 candidate rows most particles overflow their list and take the
    cell-window walk; the steps stay exact against the oracle."""
    case = _stirred(viscosity=int(general))
    nt = case["n_real"]
    print("candidate rows capped at 8, general", general, flush=True)
    orc = oracle.OracleSim(case)
    orc.run(30)
    monkeypatch.setenv("SPHGPU_VERLET_CAP", "8")
    with Simulation(case) as sim:
        sim.advance(30)
        for k in STATE_FIELDS:
            assert_close(k, sim.get(k)[:nt], orc[k][:nt], 1e-10)


@pytest.mark.parametrize("precision", [C.FP64, C.MIXED])
@pytest.mark.parametrize("which", ["stirred", "gate"])
def test_general_advance_in_pieces_equals_one_call(which, precision):
    case = _stirred() if which == "stirred" else CASES["gate"][0]()
    cfg = case["config"].copy()
    cfg.precision = precision
    print("general path in pieces:", which, precision, flush=True)
    with Simulation(case, cfg) as a, Simulation(case, cfg) as b:
        a.advance(60)
        for k in (7, 1, 25, 27):
            b.advance(k)
        for k in STATE_FIELDS:
            assert np.array_equal(a.get(k), b.get(k)), k


@pytest.mark.parametrize("precision", [C.FP64, C.MIXED])
@pytest.mark.parametrize("which", ["stirred", "gate"])
def test_general_graph_replays_equal_eager_launches(which, precision, monkeypatch):
    case = _stirred(seed=21) if which == "stirred" else CASES["gate"][0]()
    cfg = case["config"].copy()
    cfg.precision = precision
    print("general path graph vs eager:", which, precision, flush=True)
    with Simulation(case, cfg) as g:
        monkeypatch.setenv("SPHGPU_NO_GRAPH", "1")
        with Simulation(case, cfg) as e:
            for sim in (g, e):
                sim.advance(7)
            for rep in range(3):
                for sim in (g, e):
                    sim.advance(34)          # 3 graph replays + 4 eager steps
                for k in STATE_FIELDS + ("stress",) * (which == "gate"):
                    assert np.array_equal(g.get(k), e.get(k)), (rep, k)


def _mid_gate():
    """The elastic gate past the small-run size (291,750 real particles): one
    lane per fluid target, the solid targets on the side stream with several
    lanes each, and in mixed precision the filter inside the strain pass."""
    case = cases.dam_break_gate_3d(dx_f=0.0025, width=0.3)
    assert case["n_real"] > 262_144
    return case


def test_large_general_run_vs_oracle():
    case = _mid_gate()
    nt = case["n_real"]
    print("large general run: step-1 derivatives and 10 steps against the oracle", flush=True)
    orc = oracle.OracleSim(case)
    with Simulation(case) as sim:
        sim.derivatives()
        orc.derivatives()
        for k in ("dvdt", "drhodt", "dedt", "dsdt"):
            assert_close("d1 " + k, sim.get(k)[:nt], orc[k][:nt], 1e-10)
    orc = oracle.OracleSim(case)
    orc.run(10)
    pert = dict(case)
    pert["v"] = case["v"] + 1e-15 * np.random.default_rng(3).standard_normal(case["v"].shape)
    o2 = oracle.OracleSim(pert)
    o2.run(10)
    with Simulation(case) as sim:
        sim.advance(10)
        for k in STATE_FIELDS + ("stress",):
            tol = max(1e-9, 10 * rel_err(o2[k][:nt], orc[k][:nt]))
            assert_close(k, sim.get(k)[:nt], orc[k][:nt], tol)


def test_large_general_run_fused_filter_equals_separate(monkeypatch):
    """Mixed precision: the filter inside the strain pass writes the same
    exact lists in the same order as the filter pass of its own, so every
    field is bit-identical over 40 steps (rebuilds included)."""
    case = _mid_gate()
    cfg = case["config"].copy()
    cfg.precision = C.MIXED
    print("large general run, mixed: fused vs separate filter", flush=True)
    with Simulation(case, cfg) as a:
        a.advance(40)
        ra = {k: a.get(k) for k in STATE_FIELDS + ("stress",)}
    monkeypatch.setenv("SPHGPU_SEPARATE_FILTER", "1")
    with Simulation(case, cfg) as b:
        b.advance(40)
        for k in STATE_FIELDS + ("stress",):
            assert np.array_equal(b.get(k), ra[k]), k
