"""The Fortran host program on the GPU: it drives libsphgpu through its
iso_c_binding module, and its results equal those of the Python host on the
same case bit for bit (same library, same inputs, deterministic kernels).
The reference dam break read from a particle file, with the constants of
the Fortran nick preset, matches the oracle."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from mysph_amd import cases
from mysph_amd import config as C
from mysph_amd.sphgpu import Simulation
from tests.fortran_host import read_dat, run_host, write_particle_file, write_toml
from tests.parity import assert_close

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "rho", "p", "e")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_host_equals_python_host(precision, tmp_path):
    write_toml(tmp_path, f"""
name = "2D dam break"
nick = "dam_break_2d"
[Parameter]
DIM = 2
maxTimeStep = 60
saveInterval = 25
printInterval = 20
[Case]
kind = "dam_break"
dx = 0.04
precision = "{precision}"
""")
    print(f"mysph_host 2D dx=0.04 {precision} 60 steps", flush=True)
    r = run_host(tmp_path, "--final", tmp_path / "final.bin")
    assert "step 60 / 60" in r.stdout
    f = oracle.read_dump(str(tmp_path / "final.bin"))
    case = cases.dam_break_2d(dx=0.04)
    cfg = case["config"].copy()
    cfg.precision = C.FP64 if precision == "fp64" else C.MIXED
    with Simulation(case, cfg) as sim:
        sim.advance(60)
        for k in FIELDS:
            assert np.array_equal(f[k], sim.get(k).ravel()), k
    # outputs at steps 0, 25, 50 in the reference layout; step 50 matches a 50-step run
    for step in (0, 25, 50):
        assert (tmp_path / "output" / f"Type_2_{step}.dat").exists()
        assert (tmp_path / "vtk" / f"Type_-2_{step}.vtk").exists()
    names, a = read_dat(tmp_path / "output" / "Type_2_50.dat")
    with Simulation(case, cfg) as sim:
        sim.advance(50)
        nf = case["n_real"]
        np.testing.assert_allclose(a[:, 3:5], sim.get("x")[:nf], rtol=6e-8, atol=1e-12)
        np.testing.assert_allclose(a[:, 8], sim.get("rho")[:nf], rtol=6e-8)


def test_host_reference_dam_break_from_file(tmp_path):
    case = cases.reference_dam_break()
    nt = case["n_real"]
    write_particle_file(tmp_path / "ref.bin", case)
    write_toml(tmp_path, """
name = "Two-dimensional dam break"
nick = "dam_break"
[Parameter]
DIM = 2
SKF = 1
maxTimeStep = 10
saveInterval = 0
printInterval = 5
deltaT = 1e-3
summationDensity = false
dummyParticle = true
viscosity = true
gravity = true
artificialViscosity = true
[Case]
kind = "file"
file = "ref.bin"
""")
    print("mysph_host reference dam break from file, 10 steps", flush=True)
    run_host(tmp_path, "--final", tmp_path / "final.bin", "--no-output")
    f = oracle.read_dump(str(tmp_path / "final.bin"))
    orc = oracle.OracleSim(case)
    orc.run(10)
    for k in FIELDS:
        comps = 2 if k in ("x", "v") else 1
        assert_close(f"host {k}", f[k][: nt * comps], orc[k][:nt].ravel(), 1e-10)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_host_checkpoint_resume_equals_uninterrupted_run(precision, tmp_path):
    """--checkpoint after 35 steps, --resume to 80, against one 80-step run:
    the final states are bit-identical (sphgpu_checkpoint_save / _restore
    through the Fortran binding)."""
    def toml(steps):
        write_toml(tmp_path, f"""
name = "2D dam break"
nick = "dam_break_2d"
[Parameter]
DIM = 2
maxTimeStep = {steps}
saveInterval = 0
printInterval = 0
[Case]
kind = "dam_break"
dx = 0.04
precision = "{precision}"
""")
    print(f"mysph_host checkpoint/resume {precision}", flush=True)
    toml(80)
    run_host(tmp_path, "--no-output", "--final", tmp_path / "straight.bin")
    toml(35)
    run_host(tmp_path, "--no-output", "--checkpoint", tmp_path / "ck.bin")
    toml(80)
    r = run_host(tmp_path, "--no-output", "--resume", tmp_path / "ck.bin", "--final", tmp_path / "resumed.bin")
    assert "after step 35" in r.stdout
    a = oracle.read_dump(str(tmp_path / "straight.bin"))
    b = oracle.read_dump(str(tmp_path / "resumed.bin"))
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("nick,nsteps,tol", [("dam_break", 10, 1e-10), ("can_beam", 20, 1e-10)])
def test_reference_toml_runs_like_the_reference(nick, nsteps, tol, tmp_path):
    """Given only a reference sph.toml (no [Case] table), the host builds the
    reference's case and switches and its run matches the compiled
    reference's leap_frog on the same case (oracle/_ref/ref_driver <case>
    run), norm-relative per field."""
    import os
    import subprocess
    from tests.test_fortran_host import REF_TOMLS
    if not os.path.exists(oracle.REF_DRIVER):
        pytest.skip("oracle/_ref/ref_driver not built")
    toml = REF_TOMLS[nick].replace("maxTimeStep = 20000", f"maxTimeStep = {nsteps}") \
                          .replace("maxTimeStep = 100000", f"maxTimeStep = {nsteps}")
    write_toml(tmp_path, toml)
    print(f"mysph_host reference {nick}, {nsteps} steps", flush=True)
    run_host(tmp_path, "--final", tmp_path / "final.bin", "--no-output")
    f = oracle.read_dump(str(tmp_path / "final.bin"))
    out = tmp_path / "ref.bin"
    subprocess.run([oracle.REF_DRIVER, nick, "run", str(nsteps), str(out)], check=True, cwd=str(tmp_path),
                   capture_output=True, env=dict(os.environ, REF_THREADS="8"))
    ref = oracle.read_dump(str(out))
    nt, dim = int(ref["ntotal"][0]), int(ref["dim"][0])
    assert int(f["ntotal"][0]) == nt
    for k in FIELDS:
        comp = dim if k in ("x", "v") else 1
        assert_close(f"{nick} {k}", f[k][:nt * comp], ref[k][:nt * comp], tol)
