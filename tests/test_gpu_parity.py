"""GPU parity: libsphgpu (through the C ABI) against the CPU oracle on the same
inputs. The oracle itself is pinned to the compiled reference by
tests/test_oracle_golden.py.

Bars: neighbour sets bit-exact (sorted pair sets); FP64 fields within 1e-10
(norm-relative) after one step, and within the stated drift bounds after many.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from mysph_amd import cases
from mysph_amd import config as C
from mysph_amd.sphgpu import Simulation, SphGpuError
from tests.parity import assert_close, rel_err, sorted_pairs

pytestmark = pytest.mark.gpu

STATE_FIELDS = ("x", "v", "rho", "p", "e")


def oracle_pairs(case):
    off, nbr, _, _, _ = oracle.bggs(case["dim"], case["config"].skf, case["n_real"], case["x"],
                                    case["h"], case.get("state"))
    return sorted_pairs(off, nbr)


def gpu_pairs(sim):
    off, nbr = sim.pairs()
    return sorted_pairs(off, nbr)


def compare_state(sim, orc, nt, tol, fields=STATE_FIELDS, label=""):
    errs = {}
    for k in fields:
        g = sim.get(k)[:nt]
        o = orc[k][:nt]
        errs[k] = assert_close(f"{label}{k}", g, o, tol)
    return errs


# ------------------------------------------------------------- neighbour sets

def test_ref_dam_break_pairs_exact():
    case = cases.reference_dam_break()
    with Simulation(case) as sim:
        g = gpu_pairs(sim)
    o = oracle_pairs(case)
    assert len(g) == len(o) == 122267
    assert np.array_equal(g, o)


def test_ref_dam_break_grid_matches_reference():
    case = cases.reference_dam_break()
    with Simulation(case) as sim:
        sim.pairs()
        grid = sim.grid_info()
    _, _, _, _, og = oracle.bggs(2, 1, case["n_real"], case["x"], case["h"])
    assert np.array_equal(grid[:2], og[:2]) and np.array_equal(grid[3:5], og[3:5])
    assert np.array_equal(grid[6:8], og[6:8])


@pytest.mark.parametrize("k", range(6))
def test_custom_bggs_sets_exact(k):
    from tests.golden.make_golden import custom_bggs_inputs
    s = custom_bggs_inputs()[k]
    n = len(s["h"])
    case = dict(dim=s["dim"], n_real=s["ntotal"], n_dummy=n - s["ntotal"], x=s["x"],
                h=s["h"], state=s["state"], mass=np.ones(n), rho=np.full(n, 1000.0),
                type=np.where(np.arange(n) < s["ntotal"], 2, -2).astype(np.int32))
    cfg = C.base_config(s["dim"])
    cfg.skf = s["skf"]
    cfg.add_material(2, C.EOS_TAIT, 1000.0, C.tait_B(10.0), 7.0)
    case["config"] = cfg
    with Simulation(case) as sim:
        g = gpu_pairs(sim)
    o = oracle_pairs(case)
    assert np.array_equal(g, o)


# ------------------------------------------------------ reference dam break

def test_ref_dam_break_step1_derivatives():
    case = cases.reference_dam_break()
    nt = case["n_real"]
    orc = oracle.OracleSim(case)
    orc.derivatives()
    with Simulation(case) as sim:
        sim.derivatives()
        for k in ("dvdt", "drhodt", "dedt", "divx", "divv"):
            assert_close(k, sim.get(k)[:nt], orc[k][:nt], 1e-12)
        assert_close("p", sim.get("p"), orc["p"], 1e-13)
        assert_close("eta", sim.get("eta"), orc["eta"], 0.0)


@pytest.mark.parametrize("nsteps,tol", [(1, 1e-10), (10, 1e-10), (100, 1e-8)])
def test_ref_dam_break_steps(nsteps, tol):
    case = cases.reference_dam_break()
    nt = case["n_real"]
    orc = oracle.OracleSim(case)
    orc.run(nsteps)
    with Simulation(case) as sim:
        sim.advance(nsteps)
        assert sim.step == nsteps + 1
        compare_state(sim, orc, nt, tol)


def test_run_twice_bitwise_identical():
    case = cases.reference_dam_break()
    out = []
    for _ in range(2):
        with Simulation(case) as sim:
            sim.advance(20)
            out.append(np.concatenate([sim.get(k).ravel() for k in STATE_FIELDS]))
    assert np.array_equal(out[0].view(np.uint64), out[1].view(np.uint64))


def test_advance_in_pieces_equals_one_call():
    case = cases.reference_dam_break()
    with Simulation(case) as a, Simulation(case) as b:
        a.advance(12)
        b.advance(5)
        b.advance(7)
        for k in STATE_FIELDS:
            assert np.array_equal(a.get(k), b.get(k))


# ----------------------------------------------------------------- solids

@pytest.mark.parametrize("which,asw", [("2d", 0), ("2d", 1), ("3d", 0)])
def test_reference_beams(which, asw):
    from tests.golden.make_golden import beam_velocity
    case = cases.reference_can_beam() if which == "2d" else cases.reference_can_beam_3d()
    case["v"] = beam_velocity(case["x"])
    cfg = case["config"].copy()
    cfg.artificial_stress = asw
    nt = case["n_real"]
    orc = oracle.OracleSim(case, cfg)
    with Simulation(case, cfg) as sim:
        sim.derivatives()
        orc.derivatives()
        for k in ("dvdt", "drhodt", "dedt", "dsdt"):
            assert_close("d1 " + k, sim.get(k)[:nt], orc[k][:nt], 1e-10)
    orc = oracle.OracleSim(case, cfg)
    orc.run(20)
    # The reference's artificial-stress rotation (theta = atan(2 s_xy /
    # (s_xx - s_yy)) / 2, artificialTerm.f90:46-51) amplifies last-bit
    # differences: the bound is the oracle's own response to a 1e-15
    # perturbation of the input velocities (x10), or 1e-9 if that is smaller.
    tol = {k: 1e-9 for k in STATE_FIELDS + ("stress",)}
    if asw:
        pert = dict(case)
        pert["v"] = case["v"] * (1 + 1e-15 * np.random.default_rng(1).standard_normal(case["v"].shape))
        o2 = oracle.OracleSim(pert, cfg)
        o2.run(20)
        for k in tol:
            tol[k] = max(1e-9, 10 * rel_err(o2[k][:nt], orc[k][:nt]))
    with Simulation(case, cfg) as sim:
        sim.advance(20)
        for k in tol:
            assert_close(k, sim.get(k)[:nt], orc[k][:nt], tol[k])


@pytest.mark.parametrize("sph", [1, 3])
def test_sph_forms_1_and_3(sph):
    """internal_force SPH = 1 and SPH = 3 (PVRS Riemann, with the Riemann
    density rate) on the GPU against the oracle, in the reference dam break
    with its dummies at the water's sound speed."""
    c = cases.reference_dam_break()
    c["c"][c["n_real"]:] = 50.0
    cfg = c["config"].copy()
    cfg.pa_sph = sph
    nt = c["n_real"]
    orc = oracle.OracleSim(c, cfg)
    with Simulation(c, cfg) as sim:
        sim.derivatives()
        orc.derivatives()
        for k in ("dvdt", "drhodt", "dedt"):
            assert_close(f"sph {sph} d1 {k}", sim.get(k)[:nt], orc[k][:nt], 1e-10)
    orc = oracle.OracleSim(c, cfg)
    with Simulation(c, cfg) as sim:
        for n, tol in ((10, 1e-10), (90, 1e-8)):
            orc.run(n)
            sim.advance(n)
            for k in STATE_FIELDS:
                assert_close(f"sph {sph} step {sim.step - 1} {k}", sim.get(k)[:nt], orc[k][:nt], tol)


@pytest.mark.parametrize("sle", [1, 2])
def test_smoothing_length_evolution(sle):
    """h_upgrade (hsml.f90:8-38) on the GPU against the oracle (itself
    bit-exact to the reference here): the reference dam break with sle on."""
    c = cases.reference_dam_break()
    cfg = c["config"].copy()
    cfg.sle = sle
    nt = c["n_real"]
    orc = oracle.OracleSim(c, cfg)
    with Simulation(c, cfg) as sim:
        for n, tol in ((1, 1e-10), (9, 1e-10), (90, 1e-8)):
            orc.run(n)
            sim.advance(n)
            for k in STATE_FIELDS + ("h",):
                assert_close(f"sle {sle} step {sim.step - 1} {k}", sim.get(k)[:nt], orc[k][:nt], tol)


# ------------------------------------------------- benchmark dam breaks

def test_dam_break_2d_coarse_100_steps():
    case = cases.dam_break_2d(dx=0.04)
    nt = case["n_real"]
    orc = oracle.OracleSim(case)
    orc.run(100)
    with Simulation(case) as sim:
        assert np.array_equal(gpu_pairs(sim), oracle_pairs(case))
        sim.advance(100)
        compare_state(sim, orc, nt, 1e-9)


def test_dam_break_3d_coarse_20_steps():
    case = cases.dam_break_3d(dx=0.03)
    nt = case["n_real"]
    orc = oracle.OracleSim(case)
    orc.run(20)
    with Simulation(case) as sim:
        assert np.array_equal(gpu_pairs(sim), oracle_pairs(case))
        sim.advance(20)
        compare_state(sim, orc, nt, 1e-10)


# ------------------------------------------------------------- error paths

def test_nonfinite_position_rejected():
    case = cases.dam_break_2d(dx=0.1)
    case["x"] = case["x"].copy()
    case["x"][3, 0] = np.nan
    with pytest.raises(SphGpuError) as ei:
        Simulation(case)
    assert ei.value.code == -2


def test_kgc_with_fluid_targets_rejected():
    case = cases.dam_break_2d(dx=0.1)
    cfg = case["config"].copy()
    cfg.kernel_correction = 1
    with pytest.raises(SphGpuError) as ei:
        Simulation(case, cfg)
    assert ei.value.code == -4


# ------------------------------------------------------- mixed precision

def test_mixed_neighbour_sets_identical():
    case = cases.dam_break_3d(dx=0.03)
    cfg = case["config"].copy()
    cfg.precision = C.MIXED
    with Simulation(case, cfg) as sim:
        assert np.array_equal(gpu_pairs(sim), oracle_pairs(case))


def test_mixed_vs_fp64_dam_break_3d():
    """Mixed (FP32 pair physics, FP32 rho - rho0, FP64 x/v) against FP64 after
    200 steps of the coarse 3D dam break. Bounds: density 1e-5 relative,
    positions 1e-6 of the domain, velocity 1e-3 of the peak speed."""
    case = cases.dam_break_3d(dx=0.03)
    nt = case["n_real"]
    cfg = case["config"].copy()
    cfg.precision = C.MIXED
    with Simulation(case) as a, Simulation(case, cfg) as b:
        a.advance(200)
        b.advance(200)
        e_rho = rel_err(b.get("rho")[:nt], a.get("rho")[:nt])
        e_x = rel_err(b.get("x")[:nt], a.get("x")[:nt])
        e_v = rel_err(b.get("v")[:nt], a.get("v")[:nt])
        print(f"mixed vs fp64 after 200 steps: rho {e_rho:.3e} x {e_x:.3e} v {e_v:.3e}")
        assert e_rho < 1e-5 and e_x < 1e-6 and e_v < 1e-3


def test_general_and_fused_force_kernels_agree(monkeypatch):
    case = cases.dam_break_2d(dx=0.04)
    nt = case["n_real"]
    with Simulation(case) as a:
        a.advance(20)
        ra = {k: a.get(k) for k in STATE_FIELDS}
    monkeypatch.setenv("SPHGPU_FORCE_GENERAL", "1")
    with Simulation(case) as b:
        b.advance(20)
        for k in STATE_FIELDS:
            assert_close("general " + k, b.get(k)[:nt], ra[k][:nt], 1e-12)


# ------------------------------------------- candidate-list (Verlet) fast path

def test_fast_path_rebuilds_keep_neighbour_sets_exact():
    """Random velocities of ~1.5 m/s move particles across the skin many times
    in 60 steps; every rebuild must keep the candidate lists complete, so the
    trajectory still matches the oracle."""
    case = cases.dam_break_2d(dx=0.04)
    nt = case["n_real"]
    rng = np.random.default_rng(11)
    case["v"] = np.zeros_like(case["x"])
    case["v"][:nt] = rng.normal(0.0, 1.5, (nt, 2))
    orc = oracle.OracleSim(case)
    orc.run(60)
    with Simulation(case) as sim:
        sim.advance(60)
        compare_state(sim, orc, nt, 1e-8)
        st = sim.neighbour_stats()
        off, _, _, _, _ = oracle.bggs(2, 1, nt, orc["x"], case["h"])
    assert st["pairs"] > 0


@pytest.mark.parametrize("precision", [C.FP64, C.MIXED])
def test_fast_path_advance_in_pieces_equals_one_call(precision):
    """The candidate list persists across sphgpu_advance calls and is rebuilt
    only on the displacement vote, so splitting a run into calls (an output
    cadence) leaves every bit of the result unchanged, rebuilds included."""
    case = cases.dam_break_2d(dx=0.04)
    nt = case["n_real"]
    rng = np.random.default_rng(5)
    case["v"] = np.zeros_like(case["x"])
    case["v"][:nt] = rng.normal(0.0, 1.0, (nt, 2))
    cfg = case["config"].copy()
    cfg.precision = precision
    with Simulation(case, cfg) as a, Simulation(case, cfg) as b:
        a.advance(60)
        for k in (7, 1, 25, 27):
            b.advance(k)
        for k in STATE_FIELDS:
            assert np.array_equal(a.get(k), b.get(k)), k


@pytest.mark.parametrize("precision", [C.FP64, C.MIXED])
def test_graph_replays_equal_eager_launches(precision, monkeypatch):
    """Groups of steps replay a captured hipGraph; every replay, with
    downloads and synchronisation in between, equals eager launches bit for
    bit (rebuilds of the candidate list included)."""
    case = cases.dam_break_2d(dx=0.04)
    nt = case["n_real"]
    rng = np.random.default_rng(21)
    case["v"] = np.zeros_like(case["x"])
    case["v"][:nt] = rng.normal(0.0, 1.0, (nt, 2))
    cfg = case["config"].copy()
    cfg.precision = precision
    with Simulation(case, cfg) as g:
        monkeypatch.setenv("SPHGPU_NO_GRAPH", "1")
        with Simulation(case, cfg) as e:
            for sim in (g, e):
                sim.advance(7)
            for rep in range(3):
                for sim in (g, e):
                    sim.advance(34)          # 3 graph replays + 4 eager steps
                for k in STATE_FIELDS:
                    assert np.array_equal(g.get(k), e.get(k)), (rep, k)


def test_exact_list_path_matches_fast_path(monkeypatch):
    case = cases.dam_break_3d(dx=0.03)
    nt = case["n_real"]
    with Simulation(case) as a:
        a.advance(30)
        ra = {k: a.get(k) for k in STATE_FIELDS}
    monkeypatch.setenv("SPHGPU_NO_VERLET", "1")
    with Simulation(case) as b:
        b.advance(30)
        for k in STATE_FIELDS:
            assert_close("exact-list " + k, b.get(k)[:nt], ra[k][:nt], 1e-12)


# ------------------------------------------------------------ long runs

def test_ref_dam_break_1000_steps_within_1e6():
    """FP64 GPU vs the oracle (bit-identical to the reference) after 1000 steps
    of the reference dam break: within 1e-6 (norm-relative) per field."""
    case = cases.reference_dam_break()
    nt = case["n_real"]
    oracle.set_threads(0)
    orc = oracle.OracleSim(case)
    orc.run(1000)
    with Simulation(case) as sim:
        sim.advance(1000)
        errs = compare_state(sim, orc, nt, 1e-6)
        print("REF 1000 steps:", {k: f"{v:.2e}" for k, v in errs.items()})


def test_dam_break_3d_1000_steps_fp64_and_mixed():
    """3D dam break (dx = 0.03, 16,000 fluid) for 1000 steps: FP64 GPU within
    1e-6 of the oracle; mixed-precision density within 1e-4 of the oracle."""
    case = cases.dam_break_3d(dx=0.03)
    nt = case["n_real"]
    orc = oracle.OracleSim(case)
    orc.run(1000)
    with Simulation(case) as sim:
        sim.advance(1000)
        errs = compare_state(sim, orc, nt, 1e-6)
        print("3D FP64 1000 steps:", {k: f"{v:.2e}" for k, v in errs.items()})
    cfg = case["config"].copy()
    cfg.precision = C.MIXED
    with Simulation(case, cfg) as sim:
        sim.advance(1000)
        rho = sim.get("rho")[:nt]
        e_rho = float(np.max(np.abs(rho - orc["rho"][:nt]) / orc["rho"][:nt]))
        e_x = rel_err(sim.get("x")[:nt], orc["x"][:nt])
        print(f"3D mixed 1000 steps: rho {e_rho:.2e} x {e_x:.2e}")
        assert e_rho <= 1e-4


@pytest.mark.parametrize("precision", [C.FP64, C.MIXED])
@pytest.mark.parametrize("dim", [2, 3])
def test_exact_test_at_rounding_boundary(dim, precision):
    """Pairs whose acceptance a fused multiply-add in the squared distance
    would flip (tests.parity.rounding_boundary_set): the exported exact list
    and the fused force pass keep exactly the oracle's pairs."""
    from tests.parity import rounding_boundary_set
    x, h, accept = rounding_boundary_set(dim)
    n = len(h)
    case = dict(dim=dim, n_real=n, n_dummy=0, x=x, h=h, mass=np.ones(n), rho=np.full(n, 1000.0),
                type=np.full(n, 2, dtype=np.int32))
    cfg = C.base_config(dim)
    cfg.add_material(2, C.EOS_TAIT, 1000.0, C.tait_B(10.0), 7.0)
    cfg.precision = precision
    case["config"] = cfg
    o = oracle_pairs(case)
    assert o.size == 2 * int(accept.sum())
    print(f"rounding-boundary pairs dim {dim} precision {precision}: {accept.sum()} of {len(accept)} kept",
          flush=True)
    with Simulation(case, cfg) as sim:
        assert np.array_equal(gpu_pairs(sim), o)
        sim.derivatives()
        assert sim.neighbour_stats()["pairs"] == o.size


@pytest.mark.parametrize("dim", [2, 3])
def test_search_grid_at_rounding_boundary(dim):
    """A point cloud whose reference grid bounds a fused multiply-add in
    max + m (max - min) would change (tests.parity.grid_rounding_cloud): the
    device's grid equals the oracle's bit for bit."""
    from tests.parity import grid_rounding_cloud
    x = grid_rounding_cloud(dim)
    n = len(x)
    h = np.full(n, 0.05)
    case = dict(dim=dim, n_real=n, n_dummy=0, x=x, h=h, mass=np.ones(n), rho=np.full(n, 1000.0),
                type=np.full(n, 2, dtype=np.int32))
    cfg = C.base_config(dim)
    cfg.add_material(2, C.EOS_TAIT, 1000.0, C.tait_B(10.0), 7.0)
    case["config"] = cfg
    _, _, _, _, og = oracle.bggs(dim, 1, n, x, h)
    with Simulation(case, cfg) as sim:
        assert np.array_equal(gpu_pairs(sim), oracle_pairs(case))
        assert np.array_equal(sim.grid_info(), og)
