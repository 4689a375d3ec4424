"""The 2D elastic cantilever (0.2 x 0.02 m, dx = 0.5 mm, 16,120 particles,
three clamped columns, E = 2 GPa, nu = 0.3975, artificial stress eps = 0.3,
first bending mode at 0.01 m/s): the solid general path on the GPU against
the oracle, in FP64 and in mixed precision.

The reference cannot run this case (its solid constants are fixed by particle
type, eos.f90:219-248 and viscosity.f90:14-33); the oracle is pinned to the
reference on the reference's own beams (test_oracle_golden.py).

The artificial-stress rotation (theta = atan(2 s_xy / (s_xx - s_yy)) / 2,
artificialTerm.f90:46-51) switches the principal axes where s_xx - s_yy
changes sign, so the run amplifies last-bit changes: over 20,000 steps a
1e-15 relative perturbation of the initial velocities moves the oracle's
stress field by 34 % and its tip deflection by 1.6e-4. The fixture records
that response, and a FP64 result is held to max(1e-9, 10 x response).

Mixed precision rounds every pair term to FP32 (~6e-8). Bounds: after 1,000
steps every field within max(1e-5, 10 x the FP64 run's own response to a
1e-15 perturbation of the initial velocities) of FP64 (the amplification
has started by then: a change of summation order alone moves the FP64
pressure by ~2.5e-4); after 20,000 steps the tip deflection within 1e-3,
positions within 1e-5 and density within 1e-3, with the pointwise and root
bending stress reported.
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

import oracle
from mysph_amd import cases
from mysph_amd import config as C
from mysph_amd.sphgpu import Simulation
from tests.parity import assert_close, rel_err

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
FIELDS = ("x", "v", "rho", "p", "stress")


def _tip_mask(case):
    x = case["x"][:, 0]
    return x > x.max() - 0.5 * (x.max() - np.sort(np.unique(x))[-2])


def _root_moment(case, stress):
    """Bending stress at the clamp: mean s_xx of the first free column's
    upper half minus its lower half."""
    x, y = case["x"][:, 0], case["x"][:, 1]
    col = np.isclose(x, np.sort(np.unique(x[x > 0]))[0])
    sxx = stress[:, 0, 0] if stress.ndim == 3 else stress[:, 0]
    return float(np.mean(sxx[col & (y > 0)]) - np.mean(sxx[col & (y < 0)]))


def _run(case, precision, steps, every, snap=()):
    """Tip deflection every `every` steps; fields at the end and at `snap`."""
    cfg = case["config"].copy()
    cfg.precision = precision
    tip = _tip_mask(case)
    y0 = case["x"][tip, 1]
    series, snaps = [], {}
    with Simulation(case, cfg) as sim:
        for k in range(steps // every):
            sim.advance(every)
            series.append(float(np.mean(sim.get("x")[tip, 1] - y0)))
            if (k + 1) * every in snap:
                snaps[(k + 1) * every] = {f: sim.get(f) for f in FIELDS}
        snaps[steps] = {f: sim.get(f) for f in FIELDS}
    return np.array(series), snaps


def test_elastic_beam_step1_and_500_steps_vs_oracle():
    case = cases.elastic_beam()
    nt = case["n_real"]
    orc = oracle.OracleSim(case)
    with Simulation(case) as sim:
        sim.derivatives()
        orc.derivatives()
        for k in ("dvdt", "drhodt", "dedt", "dsdt"):
            assert_close("d1 " + k, sim.get(k)[:nt], orc[k][:nt], 1e-10)
    orc = oracle.OracleSim(case)
    orc.run(500)
    pert = dict(case)
    pert["v"] = case["v"] * (1 + 1e-15 * np.random.default_rng(7).standard_normal(case["v"].shape))
    o2 = oracle.OracleSim(pert)
    o2.run(500)
    with Simulation(case) as sim:
        sim.advance(500)
        for k in FIELDS:
            tol = max(1e-9, 10 * rel_err(o2[k][:nt], orc[k][:nt]))
            assert_close(k, sim.get(k)[:nt], orc[k][:nt], tol)


def test_elastic_beam_20k_steps_fp64_vs_golden():
    with open(os.path.join(GOLDEN, "elastic_beam_20k.json")) as f:
        meta = json.load(f)
    gold = np.load(os.path.join(GOLDEN, "elastic_beam_20k.npz"))
    case = cases.elastic_beam()
    assert case["n_real"] == meta["n_particles"]
    sens = meta["sensitivity_1e-15"]
    series, snaps = _run(case, C.FP64, meta["steps"], meta["tip_every"])
    out = snaps[meta["steps"]]
    print("tip deflection GPU %.9e oracle %.9e" % (series[-1], gold["tip"][-1]))
    assert_close("tip", series, gold["tip"], max(1e-9, 10 * sens["tip"]))
    nt, k = case["n_real"], meta["sample_every"]
    for f in FIELDS:
        assert_close(f, out[f][:nt:k], gold[f], max(1e-9, 10 * sens[f]))


def test_elastic_beam_mixed_vs_fp64():
    case = cases.elastic_beam()
    nt = case["n_real"]
    s64, o64 = _run(case, C.FP64, 20000, 500, snap=(1000,))
    s32, o32 = _run(case, C.MIXED, 20000, 500, snap=(1000,))
    pert = dict(case)
    pert["v"] = case["v"] * (1 + 1e-15 * np.random.default_rng(7).standard_normal(case["v"].shape))
    _, p64 = _run(pert, C.FP64, 1000, 500)
    sens = {f: rel_err(p64[1000][f][:nt], o64[1000][f][:nt]) for f in FIELDS}
    early = {f: rel_err(o32[1000][f][:nt], o64[1000][f][:nt]) for f in FIELDS}
    print("FP64 response to a 1e-15 perturbation after 1000 steps:", {k: "%.2e" % v for k, v in sens.items()})
    print("mixed vs FP64 after 1000 steps:", {k: "%.2e" % v for k, v in early.items()})
    for f in FIELDS:
        assert early[f] <= max(1e-5, 10 * sens[f]), (f, early[f], sens[f])
    a, b = o32[20000], o64[20000]
    late = {f: rel_err(a[f][:nt], b[f][:nt]) for f in FIELDS}
    late["tip"] = rel_err(s32, s64)
    late["displacement"] = rel_err(a["x"][:nt] - case["x"], b["x"][:nt] - case["x"])
    m32, m64 = _root_moment(case, a["stress"][:nt]), _root_moment(case, b["stress"][:nt])
    late["root_bending_stress"] = abs(m32 - m64) / abs(m64)
    print("mixed vs FP64 after 20000 steps:", {k: "%.2e" % v for k, v in late.items()})
    assert late["tip"] <= 1e-3
    assert late["x"] <= 1e-5
    assert late["rho"] <= 1e-3


@pytest.mark.parametrize("which,asw", [("2d", 0), ("2d", 1), ("3d", 0)])
def test_reference_beams_mixed_vs_fp64(which, asw):
    """The reference's own beams (kernel-gradient correction on) in mixed
    precision against FP64 after 20 steps: within max(1e-5, 10 x the FP64
    run's own response to a 1e-15 perturbation of the initial velocities)
    (the artificial-stress rotation amplifies last bits, as above)."""
    from tests.golden.make_golden import beam_velocity
    case = cases.reference_can_beam() if which == "2d" else cases.reference_can_beam_3d()
    case["v"] = beam_velocity(case["x"])
    nt = case["n_real"]
    pert = dict(case)
    pert["v"] = case["v"] * (1 + 1e-15 * np.random.default_rng(1).standard_normal(case["v"].shape))
    out = {}
    for key, c, prec in (("fp64", case, C.FP64), ("mixed", case, C.MIXED), ("pert", pert, C.FP64)):
        cfg = case["config"].copy()
        cfg.artificial_stress = asw
        cfg.precision = prec
        with Simulation(c, cfg) as sim:
            sim.advance(20)
            out[key] = {f: sim.get(f)[:nt] for f in FIELDS}
    err = {f: rel_err(out["mixed"][f], out["fp64"][f]) for f in FIELDS}
    sens = {f: rel_err(out["pert"][f], out["fp64"][f]) for f in FIELDS}
    print(which, asw, "mixed vs FP64", {k: "%.2e" % v for k, v in err.items()},
          "FP64 response", {k: "%.2e" % v for k, v in sens.items()})
    for f in FIELDS:
        assert err[f] <= max(1e-5, 10 * sens[f]), (f, err[f], sens[f])
