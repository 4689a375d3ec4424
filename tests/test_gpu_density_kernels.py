"""GPU against the oracle for summation density with and without Shepard
normalisation (src/Physic/density.f90:10-80) and for the Gauss and quintic
kernels (src/Numeric/kernel.f90:73-154) carried through the whole step, on
the reference dam break. The oracle equals the compiled reference bit for bit
on these runs (tests/test_oracle_golden.py: ref_dam_break_sumdens,
ref_dam_break_skf2, ref_dam_break_skf3).

Bars as for the other FP64 parity tests: 1e-12 for one derivative
evaluation, 1e-10 after 1, 10 and 20 steps (norm-relative per field); mixed
precision (FP32 pair sums over the FP64 state of the general path) within
1e-5 of FP64.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from mysph_amd import cases
from mysph_amd import config as C
from mysph_amd.sphgpu import Simulation
from tests.parity import assert_close

pytestmark = pytest.mark.gpu

STATE = ("x", "v", "rho", "p", "e")


def _case(**kw):
    case = cases.reference_dam_break()
    cfg = case["config"].copy()
    for k, v in kw.items():
        setattr(cfg, k, v)
    case["config"] = cfg
    return case


@pytest.mark.parametrize("norm", [0, 1], ids=["plain", "shepard"])
def test_summation_density_vs_oracle(norm):
    case = _case(summation_density=1, normalize_density=norm)
    nt = case["n_real"]
    print(f"summation density, shepard={norm}", flush=True)
    orc = oracle.OracleSim(case)
    orc.derivatives()
    with Simulation(case) as sim:
        sim.derivatives()
        for k in ("rho", "p"):
            assert_close(f"step-1 {k}", sim.get(k)[:nt], orc[k][:nt], 1e-12)
        for k in ("dvdt", "dedt", "divx", "divv"):
            assert_close(f"step-1 {k}", sim.get(k)[:nt], orc[k][:nt], 1e-12)
    for n in (1, 10):
        orc = oracle.OracleSim(case)
        orc.run(n)
        with Simulation(case) as sim:
            sim.advance(n)
            for k in STATE + ("divx",):
                assert_close(f"{n} steps {k}", sim.get(k)[:nt], orc[k][:nt], 1e-10)


def test_shepard_reads_the_uploaded_divergence():
    """Shepard normalisation applies where the previous step's
    divergencePosition is below 1.5 (density.f90:25-37). A caller that
    uploads the state every step (the drop-in single_step) hands the last
    step's divergences over with sphgpu_upload_field: from the state after 3
    steps, a fresh context given those divergences matches the oracle given
    the same, and a context without them matches the oracle with zeros
    (every particle normalised), and the two differ."""
    case = _case(summation_density=1, normalize_density=1)
    nt = case["n_real"]
    with Simulation(case) as sim:
        sim.advance(3)
        mid = {k: sim.get(k) for k in ("x", "v", "rho", "p", "e")}
        divx = sim.get("divx")
    interior = divx[:nt] >= 1.5
    assert nt // 2 < np.sum(interior) < nt        # both branches occur
    c2 = dict(case, **mid)
    out = {}
    for given in (True, False):
        orc = oracle.OracleSim(c2)
        if given:
            orc["divx"][:] = divx
        orc.derivatives()
        with Simulation(c2) as sim:
            if given:
                sim.put("divx", divx)
            sim.derivatives()
            g = sim.get("rho")[:nt]
        assert_close(f"rho (divergences given: {given})", g, orc["rho"][:nt], 1e-12)
        out[given] = g
    assert np.max(np.abs(out[True] - out[False])[interior] / out[False][interior]) > 1e-6


@pytest.mark.parametrize("skf", [2, 3], ids=["gauss", "quintic"])
def test_gauss_quintic_vs_oracle(skf):
    case = _case(skf=skf)
    nt = case["n_real"]
    print(f"skf={skf}", flush=True)
    orc = oracle.OracleSim(case)
    orc.derivatives()
    with Simulation(case) as sim:
        off, nbr = sim.pairs()
        o_off, o_nbr, _, _, _ = oracle.bggs(2, skf, nt, case["x"], case["h"])
        assert np.array_equal(off, o_off)
        sim.derivatives()
        for k in ("dvdt", "drhodt", "dedt", "divx", "divv"):
            assert_close(f"step-1 {k}", sim.get(k)[:nt], orc[k][:nt], 1e-12)
    orc = oracle.OracleSim(case)
    orc.run(20)
    with Simulation(case) as sim:
        sim.advance(20)
        for k in STATE:
            assert_close(f"20 steps {k}", sim.get(k)[:nt], orc[k][:nt], 1e-10)


@pytest.mark.parametrize("skf", [2, 3], ids=["gauss", "quintic"])
def test_gauss_quintic_mixed_precision(skf):
    case = _case(skf=skf)
    nt = case["n_real"]
    out = {}
    for prec in (C.FP64, C.MIXED):
        cfg = case["config"].copy()
        cfg.precision = prec
        with Simulation(case, cfg) as sim:
            sim.advance(20)
            out[prec] = {k: sim.get(k)[:nt] for k in ("x", "rho")}
    for k in ("x", "rho"):
        assert_close(f"mixed {k}", out[C.MIXED][k], out[C.FP64][k], 1e-5)
