"""The Fortran host (mysph_amd/fortran, built to mysph_amd/bin/mysph_host) on
the CPU: its case set-up equals the Python one bit for bit, its TOML reader
and output files follow the reference's layouts, and its iso_c_binding view
of sphgpu_config has the C size. Nothing here touches a GPU."""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

import oracle
from mysph_amd import cases
from mysph_amd import config as C
from tests.fortran_host import read_dat, run_host, write_particle_file, write_toml


@pytest.fixture(scope="module", autouse=True)
def built():
    from tests.fortran_host import ensure_built
    ensure_built()


def test_abi_matches_c_layout(tmp_path):
    r = run_host(tmp_path, "--abi")
    words = r.stdout.split()
    assert words[words.index("abi") + 1] == "2"
    assert int(words[words.index("config_bytes") + 1]) == ctypes.sizeof(C.Config) == 576


@pytest.mark.parametrize("dim,dx", [(2, 0.04), (3, 0.03), (2, 0.0137)])
def test_setup_matches_python_bitwise(dim, dx, tmp_path):
    write_toml(tmp_path, f"""
name = "dam break {dim}D"  # trailing comment
[Parameter]
DIM = {dim}
[Case]
kind = "dam_break"
dx = {dx!r}
""")
    run_host(tmp_path, "--setup-only", tmp_path / "setup.bin", "--no-output")
    f = oracle.read_dump(str(tmp_path / "setup.bin"))
    c = cases.dam_break_2d(dx=dx) if dim == 2 else cases.dam_break_3d(dx=dx)
    assert int(f["ntotal"][0]) == c["n_real"] and int(f["ndummy"][0]) == c["n_dummy"]
    for k in ("x", "v", "mass", "rho", "p", "e", "c", "h", "type", "state", "boundary"):
        assert np.array_equal(f[k], np.asarray(c[k]).ravel()), k


def test_case_options_and_seed(tmp_path):
    write_toml(tmp_path, """
[Parameter]
DIM = 2
[Case]
kind = "dam_break"
dx = 0.05
column = [0.5, 1.0]
tank = [2.0, 1.5]
layers = 2
h_over_dx = 1.3
jitter = 0.02
alpha = 0.15
seed = 77
""")
    run_host(tmp_path, "--setup-only", tmp_path / "s.bin", "--no-output")
    f = oracle.read_dump(str(tmp_path / "s.bin"))
    c = cases.dam_break(2, 0.05, column=(0.5, 1.0), tank=(2.0, 1.5), layers=2, h_over_dx=1.3,
                        jitter=0.02, alpha=0.15, seed=77)
    for k in ("x", "rho", "p", "c", "h", "type"):
        assert np.array_equal(f[k], np.asarray(c[k]).ravel()), k


def test_dat_and_vtk_layout(tmp_path):
    write_toml(tmp_path, """
[Parameter]
DIM = 2
[Case]
kind = "dam_break"
dx = 0.1
""")
    run_host(tmp_path, "--setup-only", tmp_path / "s.bin")
    c = cases.dam_break_2d(dx=0.1)
    nf = c["n_real"]
    names, a = read_dat(tmp_path / "output" / "Type_2_0.dat")
    # output.f90:114-123 column set for a 2D fluid
    assert names == ["Index", "Type", "State", "X", "Y", "U", "V", "Mass", "Density", "Pressure",
                     "InternalEnergy", "SoundSpeed", "SmoothingLength", "Viscosity", "DivDistance",
                     "Displacement_X", "Displacement_Y", "StressXX", "StressXY", "StressYX", "StressYY"]
    assert a.shape == (nf, len(names))
    assert np.array_equal(a[:, 0], np.arange(1, nf + 1))
    # ES15.7: 8 significant digits
    np.testing.assert_allclose(a[:, 3:5], c["x"][:nf], rtol=6e-8)
    np.testing.assert_allclose(a[:, 8], c["rho"][:nf], rtol=6e-8)
    np.testing.assert_allclose(a[:, 9], c["p"][:nf], rtol=6e-8)
    names_b, b = read_dat(tmp_path / "output" / "Type_-2_0.dat")
    assert b.shape[0] == c["n_dummy"] and np.all(b[:, 1] == -2)
    # record layout of one line: I8 index, 2(I8), 2X, then 18 x (ES15.7, 2X);
    # the final 2X moves the position but writes nothing
    line = open(tmp_path / "output" / "Type_2_0.dat").readlines()[1].rstrip("\n")
    assert len(line) == 8 + 16 + 2 + 18 * 17 - 2
    vtk = open(tmp_path / "vtk" / "Type_2_0.vtk").read().splitlines()
    assert vtk[0].startswith("# vtk DataFile Version 3.0, ")
    assert vtk[1:4] == ["paraview_vtk_output", "ASCII", "DATASET UNSTRUCTURED_GRID"]
    assert vtk[4] == f"POINTS {nf} float"
    sections = [l for l in vtk if l.startswith(("SCALARS", "VECTORS", "TENSORS", "POINT_DATA"))]
    assert sections == [f"POINT_DATA {nf}", "SCALARS Mass float 1", "SCALARS Density float 1",
                        "SCALARS Pressure float 1", "SCALARS InternalEnergy float 1",
                        "SCALARS Type int 1", "SCALARS State int 1", "SCALARS Boundary int 1",
                        "SCALARS SmoothingLength float 1", "SCALARS DivDistance float 1",
                        "VECTORS U float", "VECTORS Displacement float", "TENSORS Stress float"]
    pts = np.array([list(map(float, l.split())) for l in vtk[5:5 + nf]])
    np.testing.assert_allclose(pts[:, :2], c["x"][:nf].astype(np.float32), rtol=1e-7)
    assert np.all(pts[:, 2] == 0.0)


def test_particle_file_case(tmp_path):
    case = cases.reference_dam_break()
    write_particle_file(tmp_path / "ref.bin", case)
    write_toml(tmp_path, """
name = "Two-dimensional dam break"
nick = "dam_break"
[Parameter]
DIM = 2
summationDensity = false
viscosity = true
gravity = true
artificialViscosity = true
[Case]
kind = "file"
file = "ref.bin"
""")
    run_host(tmp_path, "--setup-only", tmp_path / "s.bin", "--no-output")
    f = oracle.read_dump(str(tmp_path / "s.bin"))
    for k in ("x", "rho", "p", "h", "type"):
        assert np.array_equal(f[k], np.asarray(case[k]).ravel()[: f[k].size]), k


@pytest.mark.parametrize("text,needle", [
    ("[Parameter]\nDIM = 4\n", "DIM must be 2 or 3"),
    ("[Parameter]\nDIM = 2\n[Case]\nkind = \"dam_break\"\ndx = fast\n", "is not a number"),
    ("[Parameter]\nDIM = 2\n[Case]\nkind = \"torus\"\n", "Case.kind"),
    ("[Parameter\nDIM = 2\n", "section header"),
    ("nick = \"dam_break\"\n[Parameter]\nDIM = 2\n[Case]\nkind = \"file\"\nfile = \"missing.bin\"\n",
     "cannot open"),
])
def test_bad_input_stops_with_message(text, needle, tmp_path):
    write_toml(tmp_path, text)
    r = run_host(tmp_path, "--setup-only", tmp_path / "s.bin", "--no-output", check=False)
    assert r.returncode != 0
    assert needle in (r.stdout + r.stderr)


def test_missing_project_file(tmp_path):
    (tmp_path / "empty").mkdir()
    r = run_host(tmp_path / "empty", check=False)
    assert r.returncode != 0 and "sph.toml" in (r.stdout + r.stderr)



# ------------------------------------------------------------- reference cases
# A reference sph.toml (the keys of data/<case>/sph.toml) has no [Case] table:
# the host builds the reference's own case for the nick and reads every switch
# as the reference does.

REF_TOMLS = {
    "dam_break": """
name = "Two-dimensional dam break"
nick = "dam_break"
[Parameter]
DIM = 2
maxN = 25000
kPair = 40
SPH = 2
SKF = 1
saveInterval = 50
maxTimeStep = 20000
deltaT = 1e-3
summationDensity    = false
dummyParticle       = true
viscosity           = true
gravity             = true
artificialViscosity = true
""",
    "can_beam": """
name = "Two-dimensional Cantilever Beam"
nick = "can_beam"
[Parameter]
DIM = 2
maxN = 505
kPair = 60
SPH = 2
SKF = 1
saveInterval = 2000
maxTimeStep = 100000
deltaT = 2e-6
summationDensity    = false
dummyParticle       = true
viscosity           = true
gravity             = true
artificialViscosity = true
averageVelocity     = false
kernelCorrection    = true
artificialStress = false
""",
    "can_beam_3d": """
name = "Three-dimensional Cantilever Beam"
nick = "can_beam_3d"
[Parameter]
DIM = 3
maxN = 9045
kPair = 400
SPH = 2
SKF = 1
saveInterval = 2000
maxTimeStep = 100000
deltaT = 2e-6
summationDensity    = false
dummyParticle       = true
viscosity           = true
gravity             = true
artificialViscosity = true
kernelCorrection    = true
""",
}

REF_CASES = {"dam_break": cases.reference_dam_break, "can_beam": cases.reference_can_beam,
             "can_beam_3d": cases.reference_can_beam_3d}


@pytest.mark.parametrize("nick", sorted(REF_TOMLS))
def test_reference_toml_builds_the_reference_case(nick, tmp_path):
    write_toml(tmp_path, REF_TOMLS[nick])
    r = run_host(tmp_path, "--setup-only", tmp_path / "setup.bin", "--no-output")
    assert "case reference" in r.stdout
    f = oracle.read_dump(str(tmp_path / "setup.bin"))
    c = REF_CASES[nick]()
    assert int(f["ntotal"][0]) == c["n_real"] and int(f["ndummy"][0]) == c["n_dummy"]
    for k in ("x", "v", "mass", "rho", "p", "e", "c", "h", "type", "state", "boundary"):
        assert np.array_equal(f[k], np.asarray(c[k]).ravel()), k


@pytest.mark.skipif(not os.path.exists(oracle.REF_DRIVER),
                    reason="oracle/_ref/ref_driver not built (needs the reference sources at build time)")
@pytest.mark.parametrize("nick", sorted(REF_TOMLS))
def test_output_equals_the_reference_writer(nick, tmp_path):
    """The .dat and .vtk files of the set-up state, byte for byte against the
    reference's own output routine (output.f90:7-352) on the reference's own
    set-up (oracle/_ref/ref_driver <case> output); the first line of a VTK
    file carries the time of writing."""
    import subprocess
    ref = tmp_path / "ref"
    ref.mkdir()
    subprocess.run([oracle.REF_DRIVER, nick, "output", "0", str(ref / "unused.bin")], check=True,
                   cwd=str(ref), capture_output=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
    host = tmp_path / "host"
    write_toml(host, REF_TOMLS[nick])
    run_host(host, "--setup-only", host / "setup.bin")
    want = sorted(os.listdir(ref / "output"))
    assert want and sorted(os.listdir(host / "output")) == want
    assert sorted(os.listdir(host / "vtk")) == sorted(os.listdir(ref / "vtk"))
    for name in want:
        assert (host / "output" / name).read_bytes() == (ref / "output" / name).read_bytes(), name
    for name in os.listdir(ref / "vtk"):
        a = (host / "vtk" / name).read_bytes().split(b"\n", 1)
        b = (ref / "vtk" / name).read_bytes().split(b"\n", 1)
        assert a[0].startswith(b"# vtk DataFile Version 3.0") and b[0].startswith(b"# vtk DataFile")
        assert a[1] == b[1], name


def test_unknown_nick_without_case_is_refused(tmp_path):
    write_toml(tmp_path, """
name = "Water impact"
nick = "water_impact"
[Parameter]
DIM = 2
""")
    r = run_host(tmp_path, "--setup-only", tmp_path / "s.bin", "--no-output", check=False)
    assert r.returncode != 0
    assert "not a case this host builds" in r.stdout


@pytest.mark.parametrize("key", ["averageVelocity", "artificialHeat", "openBoundary", "shifting"])
def test_unsupported_switch_is_refused(key, tmp_path):
    write_toml(tmp_path, REF_TOMLS["dam_break"].replace("[Parameter]", f"[Parameter]\n{key} = true"))
    r = run_host(tmp_path, "--setup-only", tmp_path / "s.bin", "--no-output", check=False)
    assert r.returncode != 0 and key in r.stdout
