"""GPU against the oracle on the seeded sets that pin the reference's
external-force pair branches and the J2 cap (tests/golden/custom_sets.py;
the oracle equals the compiled reference bit for bit on them,
tests/test_oracle_golden.py):

  - the fluid-fluid penalty (externalForce.f90:126-141), the fluid-solid
    coupling (:142-170), the wall force at the default range (:109-124) and
    gravity with solids skipped (nicks db_gate, water_impact, :42-48), alone
    (every other force is zero: no EOS, zero pressure, no viscosity) and
    with the full physics set, on the general path and, for the fluid-only
    set, on the fused fast path;
  - leap_frog with type-104 solids whose deviatoric stress reaches the
    yield surface (timeIntegration.f90:149-166, :190-205, :225-240).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from mysph_amd import config as C
from mysph_amd.sphgpu import Simulation
from tests.golden.custom_sets import (as_case, external_force_sets, j2_config, j2_sets,
                                      nick_config)
from tests.parity import assert_close

pytestmark = pytest.mark.gpu

DT = 1e-4


def _ext_only(cfg, types):
    """No EOS (pressure stays the uploaded zero) for every type present, so
    the internal force is zero and dv/dt is the external force."""
    for t in sorted(set(abs(int(t)) for t in types)):
        cfg.add_material(t, C.EOS_NONE, 1000.0, 0.0)
    return cfg


def _physics(cfg, types):
    """Tait water for fluid types, the reference's type-102 solid."""
    for t in sorted(set(abs(int(t)) for t in types)):
        if t > 100:
            cfg.add_material(t, C.EOS_LINEAR, 1100.0, C.f32(2e7) / 1100.0, eta=C.f32(4.27e6))
        else:
            cfg.add_material(t, C.EOS_TAIT, 1000.0, C.tait_B(30.0), 7.0, eta=C.f32(1e-3))
    cfg.artificial_viscosity = 1
    return cfg


@pytest.mark.parametrize("general", [False, True], ids=["default_path", "general_path"])
@pytest.mark.parametrize("k", range(3))
def test_external_force_alone_vs_oracle(k, general, monkeypatch):
    name, nick, dim, flags, ntotal, s = external_force_sets()[k]
    print(f"set {name} general={general}", flush=True)
    if general:
        monkeypatch.setenv("SPHGPU_FORCE_GENERAL", "1")
    cfg = _ext_only(nick_config(dim, nick, flags, DT, s["x"]), s["type"])
    case = as_case(s, ntotal, dim, cfg)
    want = oracle.OracleSim(case).external_force()
    with Simulation(case) as sim:
        sim.derivatives()
        got = sim.get("dvdt")[:ntotal]
    assert_close(f"{name} dv/dt", got, want, 1e-12)
    # per particle too: forces span many orders of magnitude here
    scale = np.maximum(np.abs(want).max(axis=1), 1e-300)
    assert np.max(np.abs(got - want).max(axis=1) / scale) < 1e-10


@pytest.mark.parametrize("k", range(3))
def test_external_force_sets_full_physics_vs_oracle(k):
    name, nick, dim, flags, ntotal, s = external_force_sets()[k]
    print(f"set {name}", flush=True)
    cfg = _physics(nick_config(dim, nick, flags, DT, s["x"]), s["type"])
    case = as_case(s, ntotal, dim, cfg)
    orc = oracle.OracleSim(case)
    orc.derivatives()
    with Simulation(case) as sim:
        sim.derivatives()
        for f in ("dvdt", "drhodt", "dedt"):
            assert_close(f"{name} {f}", sim.get(f)[:ntotal], orc[f][:ntotal], 1e-12)


def test_coupling_trajectory_vs_oracle():
    """50 leap-frog steps of the 2D fluid-solid set (coupling, gravity with
    the solid skipped, Tait water, linear solid, artificial viscosity) at a
    step that resolves its forces. A 1e-15 perturbation of the oracle's own
    input moves its velocities by ~3e-14 here, so the bar is 1e-10."""
    name, nick, dim, flags, ntotal, s = external_force_sets()[1]
    cfg = _physics(nick_config(dim, nick, flags, 2e-6, s["x"]), s["type"])
    case = as_case(s, ntotal, dim, cfg)
    orc = oracle.OracleSim(case)
    orc.run(50)
    with Simulation(case) as sim:
        sim.advance(50)
        for f in ("x", "v", "rho", "p"):
            assert_close(f"{name} 50 steps {f}", sim.get(f)[:ntotal], orc[f][:ntotal], 1e-10)


def test_penalty_set_mixed_precision():
    """The fluid-only set on the fast path in mixed precision against FP64."""
    name, nick, dim, flags, ntotal, s = external_force_sets()[0]
    out = {}
    for prec in (C.FP64, C.MIXED):
        cfg = _ext_only(nick_config(dim, nick, flags, DT, s["x"]), s["type"])
        cfg.precision = prec
        with Simulation(as_case(s, ntotal, dim, cfg)) as sim:
            sim.derivatives()
            out[prec] = sim.get("dvdt")[:ntotal]
    assert_close("mixed dv/dt", out[C.MIXED], out[C.FP64], 1e-5)


@pytest.mark.parametrize("k", range(2))
def test_j2_cap_vs_oracle(k):
    name, dim, flags, nsteps, dt, s = j2_sets()[k]
    print(f"set {name}", flush=True)
    nt = len(s["h"])
    case = as_case(s, nt, dim, j2_config(dim, dt, flags))
    orc = oracle.OracleSim(case)
    orc.derivatives()
    with Simulation(case) as sim:
        sim.derivatives()
        for f in ("dvdt", "drhodt", "dedt", "dsdt"):
            assert_close(f"{name} step-1 {f}", sim.get(f)[:nt], orc[f][:nt], 1e-12)
    # the cap engages during these runs (tests/test_oracle_golden.py,
    # test_j2_cap_engages); it is checked here against the oracle's
    orc = oracle.OracleSim(case)
    orc.run(nsteps)
    ys = np.sqrt(3e8 ** 2 / 3)
    with Simulation(case) as sim:
        sim.advance(nsteps)
        for f in ("x", "v", "rho", "p"):
            assert_close(f"{name} {nsteps} steps {f}", sim.get(f)[:nt], orc[f][:nt], 1e-10)
        for f in ("shear", "stress"):
            assert_close(f"{name} {nsteps} steps {f}", sim.get(f)[:nt], orc[f][:nt], 1e-9)
        sh = sim.get("shear")[:nt]
    assert np.max(np.sqrt(np.sum(sh ** 2, axis=(1, 2)))) <= ys * (1 + 1e-12)
