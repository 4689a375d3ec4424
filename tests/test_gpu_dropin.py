"""Drop-in parity: the reference's own program (case set-up, dummy particles,
leap_frog) with its single_step replaced at link time by
mysph_amd/fortran/single_step_gpu.f90, against the same program with the
reference's single_step (src/TimeIntegration/singleStep.f90).

Both binaries are built by oracle/build_ref.sh from the reference sources;
they differ only in the object that defines single_step. Bars as for the
other FP64 parity tests: 1e-12 for one derivative evaluation, 1e-10 after
one and ten steps, 1e-8 after 100 (norm-relative per field).
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import oracle
from tests.parity import assert_close

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
GPU = os.path.join(ROOT, "oracle", "_ref", "ref_driver_gpu")

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(not (os.path.exists(REF) and os.path.exists(GPU)),
                       reason="oracle/_ref drivers not built (needs the reference sources at build time)"),
]

STATE = ("x", "v", "rho", "p", "e")


def run(exe, case, mode, nsteps, tmp_path, threads=8, extra_env=None):
    out = tmp_path / f"{os.path.basename(exe)}_{case}_{mode}_{nsteps}.bin"
    env = dict(os.environ, REF_THREADS=str(threads), **(extra_env or {}))
    print(f"run {os.path.basename(exe)} {case} {mode} {nsteps}", flush=True)
    r = subprocess.run([exe, case, mode, str(nsteps), str(out)], env=env, cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, f"{exe} failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}"
    return oracle.read_dump(str(out))


def nreal(d, k, comps=1):
    return d[k][: int(d["ntotal"][0]) * comps]


@pytest.mark.parametrize("case", ["dam_break", "can_beam", "can_beam_3d"])
def test_dropin_derivatives(case, tmp_path):
    ref = run(REF, case, "deriv", 1, tmp_path)
    gpu = run(GPU, case, "deriv", 1, tmp_path)
    dim = int(ref["dim"][0])
    assert int(gpu["ntotal"][0]) == int(ref["ntotal"][0])
    assert int(gpu["ndummy"][0]) == int(ref["ndummy"][0])
    for k in ("dvdt", "drhodt", "dedt"):
        assert_close(f"{case} {k}", gpu[k], ref[k], 1e-12)
    if case != "dam_break":
        assert_close(f"{case} dsdt", gpu["dsdt"], ref["dsdt"], 1e-12)
        assert_close(f"{case} stress", nreal(gpu, "stress", dim * dim), nreal(ref, "stress", dim * dim), 1e-12)
    assert_close(f"{case} p", gpu["p"], ref["p"], 1e-12)
    assert_close(f"{case} divv", nreal(gpu, "divv"), nreal(ref, "divv"), 1e-12)
    # the state the reference does not change is passed through untouched
    for k in ("x", "v", "mass", "h", "type"):
        assert np.array_equal(gpu[k], ref[k]), k


@pytest.mark.parametrize("case,nsteps,tol", [("dam_break", 1, 1e-10), ("dam_break", 10, 1e-10),
                                             ("dam_break", 100, 1e-8), ("can_beam", 200, 1e-8)])
def test_dropin_leap_frog(case, nsteps, tol, tmp_path):
    ref = run(REF, case, "run", nsteps, tmp_path)
    gpu = run(GPU, case, "run", nsteps, tmp_path)
    dim = int(ref["dim"][0])
    for k in STATE:
        comps = dim if k in ("x", "v") else 1
        assert_close(f"{case} {nsteps} {k}", nreal(gpu, k, comps), nreal(ref, k, comps), tol)


@pytest.mark.parametrize("sle", [1, 2])
def test_dropin_smoothing_length_evolution(sle, tmp_path):
    """Config%sle on: h_upgrade (hsml.f90:8-38) runs inside the GPU
    single_step and the updated SmoothingLength comes back."""
    env = {"REF_SLE": str(sle), "REF_PAIRNUM": "200"}
    ref = run(REF, "dam_break", "run", 10, tmp_path, extra_env=env)
    gpu = run(GPU, "dam_break", "run", 10, tmp_path, extra_env=env)
    for k in STATE + ("h",):
        comps = 2 if k in ("x", "v") else 1
        assert_close(f"sle {sle} {k}", nreal(gpu, k, comps), nreal(ref, k, comps), 1e-10)


@pytest.mark.parametrize("sph", [1, 3])
def test_dropin_sph_forms(sph, tmp_path):
    """Config%pa_sph = 1 and 3 (Riemann) through the GPU single_step; the
    dummies carry the water's sound speed (the Riemann forms divide by
    rho_i c + rho_j c)."""
    env = {"REF_PA_SPH": str(sph), "REF_DUMMY_C": "50"}
    ref = run(REF, "dam_break", "run", 10, tmp_path, extra_env=env)
    gpu = run(GPU, "dam_break", "run", 10, tmp_path, extra_env=env)
    for k in STATE:
        comps = 2 if k in ("x", "v") else 1
        assert_close(f"sph {sph} {k}", nreal(gpu, k, comps), nreal(ref, k, comps), 1e-10)


@pytest.mark.parametrize("norm", ["0", "1"], ids=["plain", "shepard"])
def test_dropin_summation_density(norm, tmp_path):
    """Config%sum_density_w on, with and without Shepard normalisation
    (density.f90:10-80): the GPU single_step writes Density back, and with
    normalisation it first uploads the previous step's divergencePosition,
    which decides which particles are normalised (density.f90:25-37)."""
    env = {"REF_SUM_DENSITY": "1", "REF_NORM_DENSITY": norm}
    for nsteps in (1, 10, 30):
        ref = run(REF, "dam_break", "run", nsteps, tmp_path, extra_env=env)
        gpu = run(GPU, "dam_break", "run", nsteps, tmp_path, extra_env=env)
        for k in STATE + ("divx",):
            comps = 2 if k in ("x", "v") else 1
            assert_close(f"sum density {norm} {nsteps} {k}", nreal(gpu, k, comps),
                         nreal(ref, k, comps), 1e-10)


@pytest.mark.parametrize("skf", ["2", "3"], ids=["gauss", "quintic"])
def test_dropin_gauss_quintic(skf, tmp_path):
    """Config%skf = 2 (Gauss) and 3 (quintic) through the GPU single_step
    (kernel.f90:73-154); the search reaches 3h (REF_PAIRNUM = 60)."""
    env = {"REF_SKF": skf, "REF_PAIRNUM": "60"}
    ref = run(REF, "dam_break", "deriv", 1, tmp_path, extra_env=env)
    gpu = run(GPU, "dam_break", "deriv", 1, tmp_path, extra_env=env)
    for k in ("dvdt", "drhodt", "dedt"):
        assert_close(f"skf {skf} {k}", gpu[k], ref[k], 1e-12)
    ref = run(REF, "dam_break", "run", 10, tmp_path, extra_env=env)
    gpu = run(GPU, "dam_break", "run", 10, tmp_path, extra_env=env)
    for k in STATE:
        comps = 2 if k in ("x", "v") else 1
        assert_close(f"skf {skf} 10 steps {k}", nreal(gpu, k, comps), nreal(ref, k, comps), 1e-10)
