"""Writes the golden fixtures in tests/golden/ from the compiled reference.

Runs oracle/_ref/ref_driver (the reference's own Fortran modules, built by
oracle/build_ref.sh from /root/reference/src) on:
  - the reference 2D dam break (input.f90:1031-1088 + dummyParticles.f90:
    544-659, data/damBreak/sph.toml): initial state, BGGS pairs, step-1
    derivatives, states after 1/10/100/1000 leap-frog steps, and the
    summation-density (+ Shepard) variant;
  - the cantilever beams (can_beam with and without artificial stress,
    can_beam_3d), with a deterministic velocity field so the solid terms are
    non-zero from step 1;
  - BGGS alone on seeded random particle sets (2D/3D, cubic/Gauss/quintic
    kernels, State == -1 particles, dummies, windows that miss neighbours).

Fixtures are data only: inputs (or the recipe to regenerate them) and the
reference's outputs, as sampled values and SHA-256 hashes of full arrays
(little-endian float64 / int32 bytes). Run in the build container:
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from mysph_amd import cases  # noqa: E402

DRIVER = oracle.REF_DRIVER
SAMPLE = 16          # every 16th real particle for the dam break
STEPS = (1, 10, 100, 1000)
STATE_H = ("x", "v", "rho", "p", "e", "h", "divv")


def sha(a) -> str:
    a = np.ascontiguousarray(a)
    if a.dtype.kind == "f":
        a = a.astype("<f8")
    else:
        a = a.astype("<i4")
    return hashlib.sha256(a.tobytes()).hexdigest()


def run(case, mode, nsteps, extra_env=None, vin=None):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "out.bin")
        cmd = [DRIVER, case, mode, str(nsteps), out]
        if vin is not None:
            vpath = os.path.join(td, "v.bin")
            np.ascontiguousarray(vin, dtype="<f8").tofile(vpath)
            cmd.append(vpath)
        env = dict(os.environ)
        env["OMP_NUM_THREADS"] = "1"
        if extra_env:
            env.update(extra_env)
        subprocess.run(cmd, check=True, env=env, cwd=td, stdout=subprocess.DEVNULL)
        return oracle.read_dump(out)


def beam_velocity(x: np.ndarray) -> np.ndarray:
    """Deterministic velocity field used for the beam fixtures."""
    L = x[:, 0].max()
    v = np.zeros_like(x)
    v[:, 1] = 0.5 * (x[:, 0] / L) ** 2 + 0.05 * np.sin(37 * x[:, 0] / L)
    if x.shape[1] == 3:
        v[:, 2] = 0.1 * (x[:, 0] / L) * np.cos(11 * x[:, 1] / L)
    return v


def state_record(d, nt, sample, fields=("x", "v", "rho", "p", "e", "disp", "divx", "divv",
                                        "stress")):
    rec = {}
    dim = int(d["dim"][0])
    for k in fields:
        a = d[k]
        comp = a.size // (int(d["ntotal"][0]) + int(d["ndummy"][0]))
        a = a.reshape(-1, comp)[:nt]
        rec[k] = a[::sample].copy()
        rec["sha_" + k] = sha(a)
    rec["dim"] = dim
    return rec


def save(name, arrays: dict, meta: dict):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **{k: v for k, v in arrays.items()})
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("wrote", path, os.path.getsize(path), "bytes")


def flat(prefix, rec):
    arrays, meta = {}, {}
    for k, v in rec.items():
        if k.startswith("sha_"):
            meta[prefix + k] = v
        elif isinstance(v, np.ndarray):
            arrays[prefix + k] = v
        else:
            meta[prefix + k] = v
    return arrays, meta


def dam_break():
    case = cases.reference_dam_break()
    nt = case["n_real"]
    arrays, meta = {}, {"sample": SAMPLE, "ntotal": nt, "ndummy": case["n_dummy"]}
    d = run("dam_break", "init", 0)
    for k in ("x", "mass", "rho", "p", "h", "c", "type"):
        meta["init_sha_" + k] = sha(d[k])
    d = run("dam_break", "pairs", 0)
    cnt = d["count"]
    meta["pairs_sha_count"] = sha(cnt)
    meta["pairs_sha_nbr"] = sha(d["nbr"])
    meta["pairs_sha_w"] = sha(d["w"])
    meta["pairs_sha_dwdx"] = sha(d["dwdx"])
    meta["pairs_total"] = int(cnt.sum())
    off = np.concatenate([[0], np.cumsum(cnt)])
    pick = np.arange(0, nt, 64)
    sel = np.concatenate([np.arange(off[i], off[i + 1]) for i in pick])
    arrays["pairs_targets"] = pick.astype(np.int32)
    arrays["pairs_count"] = cnt.astype(np.int32)
    arrays["pairs_nbr_sample"] = d["nbr"][sel].astype(np.int32)
    arrays["pairs_w_sample"] = d["w"][sel]
    arrays["pairs_dwdx_sample"] = d["dwdx"].reshape(-1, 2)[sel]
    d = run("dam_break", "deriv", 1)
    for k, comp in (("dvdt", 2), ("drhodt", 1), ("dedt", 1)):
        a = d[k].reshape(-1, comp)[:nt]
        arrays["deriv_" + k] = a[::SAMPLE].copy()
        meta["deriv_sha_" + k] = sha(a)
    for k in ("p", "divx", "divv", "eta"):
        a = d[k][:nt]
        arrays["deriv_" + k] = a[::SAMPLE].copy()
        meta["deriv_sha_" + k] = sha(a)
    meta["deriv_sha_p_all"] = sha(d["p"])
    for n in STEPS:
        d = run("dam_break", "run", n)
        a, m = flat(f"step{n}_", state_record(d, nt, SAMPLE))
        arrays.update(a)
        meta.update(m)
    save("ref_dam_break", arrays, meta)

    arrays, meta = {}, {"sample": SAMPLE}
    env = {"REF_SUM_DENSITY": "1", "REF_NORM_DENSITY": "1"}
    for n in (1, 10):
        d = run("dam_break", "run", n, env)
        a, m = flat(f"step{n}_", state_record(d, nt, SAMPLE))
        arrays.update(a)
        meta.update(m)
    save("ref_dam_break_sumdens", arrays, meta)
    dam_break_sle()


def dam_break_sph13():
    """The other pair forms of internal_force: SPH = 1 (internalForce.f90:
    197-225) and SPH = 3 with the Riemann density rate (:277-325,
    density.f90:109-145), in the reference dam break with its dummies given
    the water's sound speed (50, eos.f90:37; the Riemann forms divide by
    rho_i c + rho_j c, which is 0/0 for a dummy pair at c = 0)."""
    nt = cases.reference_dam_break()["n_real"]
    for sph in (1, 3):
        arrays, meta = {}, {"sample": SAMPLE, "pa_sph": sph, "dummy_c": 50.0}
        d = run("dam_break", "deriv", 1, {"REF_PA_SPH": str(sph), "REF_DUMMY_C": "50"})
        for k, comp in (("dvdt", 2), ("drhodt", 1), ("dedt", 1)):
            a = d[k].reshape(-1, comp)[:nt]
            arrays["deriv_" + k] = a[::SAMPLE].copy()
            meta["deriv_sha_" + k] = sha(a)
        for n in (1, 10, 100):
            d = run("dam_break", "run", n, {"REF_PA_SPH": str(sph), "REF_DUMMY_C": "50"})
            a, m = flat(f"step{n}_", state_record(d, nt, SAMPLE, fields=STATE_H))
            arrays.update(a)
            meta.update(m)
        save(f"ref_dam_break_sph{sph}", arrays, meta)


def dam_break_skf():
    """The Gauss (SKF = 2) and quintic (SKF = 3) kernels (kernel.f90:73-154)
    in the reference dam break: the search reaches 3h, so the pair list holds
    about 28 neighbours (REF_PAIRNUM = 60); step-1 derivatives and the
    state after 1 and 10 steps."""
    nt = cases.reference_dam_break()["n_real"]
    for skf in (2, 3):
        arrays, meta = {}, {"sample": SAMPLE, "skf": skf}
        env = {"REF_SKF": str(skf), "REF_PAIRNUM": "60"}
        d = run("dam_break", "deriv", 1, env)
        for k, comp in (("dvdt", 2), ("drhodt", 1), ("dedt", 1)):
            a = d[k].reshape(-1, comp)[:nt]
            arrays["deriv_" + k] = a[::SAMPLE].copy()
            meta["deriv_sha_" + k] = sha(a)
        for n in (1, 10):
            d = run("dam_break", "run", n, env)
            a, m = flat(f"step{n}_", state_record(d, nt, SAMPLE, fields=STATE_H))
            arrays.update(a)
            meta.update(m)
        save(f"ref_dam_break_skf{skf}", arrays, meta)


def dam_break_sle():
    """Smoothing-length evolution (hsml.f90:8-38) switched on in the
    reference dam break."""
    nt = cases.reference_dam_break()["n_real"]
    for sle in (1, 2):
        arrays, meta = {}, {"sample": SAMPLE, "sle": sle}
        for n in (1, 10, 100):
            # h = 2 dx after the first update: room for ~50 neighbours
            d = run("dam_break", "run", n, {"REF_SLE": str(sle), "REF_PAIRNUM": "200"})
            a, m = flat(f"step{n}_", state_record(d, nt, SAMPLE, fields=STATE_H))
            arrays.update(a)
            meta.update(m)
        save(f"ref_dam_break_sle{sle}", arrays, meta)


def beams():
    for name, fn, env, tag in (
        ("can_beam", cases.reference_can_beam, None, "ref_can_beam"),
        ("can_beam", cases.reference_can_beam, {"REF_ARTI_STRESS": "1"}, "ref_can_beam_as"),
        ("can_beam_3d", cases.reference_can_beam_3d, None, "ref_can_beam_3d"),
    ):
        case = fn()
        nt = case["n_real"]
        sample = 2 if case["dim"] == 2 else 32
        v = beam_velocity(case["x"])
        arrays, meta = {}, {"sample": sample, "ntotal": nt}
        d = run(name, "deriv", 1, env, v)
        for k, comp in (("dvdt", case["dim"]), ("drhodt", 1), ("dedt", 1),
                        ("dsdt", case["dim"] ** 2)):
            a = d[k].reshape(-1, comp)[:nt]
            arrays["deriv_" + k] = a[::sample].copy()
            meta["deriv_sha_" + k] = sha(a)
        for n in (1, 20, 200):
            d = run(name, "run", n, env, v)
            a, m = flat(f"step{n}_", state_record(d, nt, sample))
            arrays.update(a)
            meta.update(m)
        save(tag, arrays, meta)


def custom_bggs_inputs():
    """Seeded random particle sets for BGGS (regenerated by the tests)."""
    rng = np.random.default_rng(20261205)
    sets = []
    # 2D, cubic, window ~0.9 cell: int(h/cell)+1 = 1 < 2h/cell, misses pairs
    n = 300
    x = rng.random((n, 2))
    sets.append(dict(dim=2, skf=1, ntotal=250, x=x, h=np.full(n, 0.115),
                     state=np.zeros(n, np.int32)))
    # 2D, cubic, per-particle h and State == -1 particles
    n = 400
    x = rng.random((n, 2)) * [2.0, 1.0]
    h = 0.03 + 0.05 * rng.random(n)
    st = np.where(rng.random(n) < 0.05, -1, 0).astype(np.int32)
    sets.append(dict(dim=2, skf=1, ntotal=350, x=x, h=h, state=st))
    # 2D Gauss and quintic (scale_k = 3)
    n = 250
    x = rng.random((n, 2))
    sets.append(dict(dim=2, skf=2, ntotal=250, x=x, h=np.full(n, 0.04),
                     state=np.zeros(n, np.int32)))
    sets.append(dict(dim=2, skf=3, ntotal=200, x=x, h=np.full(n, 0.05),
                     state=np.zeros(n, np.int32)))
    # 3D cubic, window miss, jittered lattice with exact ties removed
    n = 500
    x = rng.random((n, 3))
    sets.append(dict(dim=3, skf=1, ntotal=450, x=x, h=np.full(n, 0.14),
                     state=np.zeros(n, np.int32)))
    # 3D lattice (exact r == 2h ties) with dummies
    g = np.stack(np.meshgrid(*[np.arange(8)] * 3, indexing="ij"), -1).reshape(-1, 3) * 0.1
    sets.append(dict(dim=3, skf=1, ntotal=400, x=g.astype(np.float64),
                     h=np.full(len(g), 0.1), state=np.zeros(len(g), np.int32)))
    return sets


def custom_bggs():
    arrays, meta = {}, {"nsets": 0}
    for k, s in enumerate(custom_bggs_inputs()):
        n = len(s["h"])
        with tempfile.TemporaryDirectory() as td:
            inp = os.path.join(td, "in.bin")
            out = os.path.join(td, "out.bin")
            with open(inp, "wb") as f:
                np.array([s["dim"], s["skf"], s["ntotal"], n, 400], "<i4").tofile(f)
                np.ascontiguousarray(s["x"], "<f8").tofile(f)
                np.ascontiguousarray(s["h"], "<f8").tofile(f)
                np.ascontiguousarray(s["state"], "<i4").tofile(f)
            subprocess.run([DRIVER, "custom", "bggs", "0", out, inp], check=True, cwd=td,
                           stdout=subprocess.DEVNULL, env={**os.environ, "OMP_NUM_THREADS": "1"})
            d = oracle.read_dump(out)
        arrays[f"set{k}_count"] = d["count"].astype(np.int32)
        arrays[f"set{k}_nbr_head"] = d["nbr"][:1000].astype(np.int32)
        arrays[f"set{k}_w_head"] = d["w"][:1000]
        arrays[f"set{k}_dwdx_head"] = d["dwdx"].reshape(-1, s["dim"])[:1000]
        meta[f"set{k}_sha_nbr"] = sha(d["nbr"])
        meta[f"set{k}_sha_w"] = sha(d["w"])
        meta[f"set{k}_sha_dwdx"] = sha(d["dwdx"])
        meta[f"set{k}_sha_x"] = sha(s["x"])
        meta["nsets"] += 1
    save("ref_bggs_custom", arrays, meta)


def run_custom(mode, nsteps, s, dim, ntotal, flags, nick, dt, kpair=200):
    """ref_driver on a custom particle state (tests/golden/custom_sets.py)."""
    from tests.golden.custom_sets import write_driver_input
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "in.bin")
        out = os.path.join(td, "out.bin")
        write_driver_input(inp, s, dim, 1, ntotal, kpair, flags, nick, dt)
        subprocess.run([DRIVER, "custom", mode, str(nsteps), out, inp], check=True, cwd=td,
                       stdout=subprocess.DEVNULL, env={**os.environ, "OMP_NUM_THREADS": "1"})
        return oracle.read_dump(out)


def ext_force():
    """external_force alone (externalForce.f90:7-174) on the custom sets:
    penalty, coupling, wall force and the per-nick gravity variants."""
    from tests.golden.custom_sets import external_force_sets
    arrays, meta = {}, {"sets": []}
    for name, nick, dim, flags, ntotal, s in external_force_sets():
        d = run_custom("extforce", 0, s, dim, ntotal, flags, nick, 1e-4)
        a = d["exdvdt"].reshape(-1, dim)
        meta["sets"].append(name)
        meta[f"{name}_nick"] = nick
        meta[f"{name}_sha_exdvdt"] = sha(a)
        meta[f"{name}_sha_count"] = sha(d["count"])
        meta[f"{name}_sha_x"] = sha(s["x"])
        meta[f"{name}_pairs"] = int(d["count"].sum())
        arrays[f"{name}_exdvdt"] = a
    save("ref_ext_force", arrays, meta)


def j2_cap():
    """leap_frog with the J2 cap of type 104 (timeIntegration.f90:149-166,
    :190-205, :225-240) on seeded blocks; state after 1, 30 and 60 steps."""
    from tests.golden.custom_sets import j2_sets
    arrays, meta = {"_": np.zeros(1)}, {"sets": []}
    for name, dim, flags, nsteps, dt, s in j2_sets():
        meta["sets"].append(name)
        meta[f"{name}_sha_x0"] = sha(s["x"])
        nt = len(s["h"])
        for n in (1, nsteps // 2, nsteps):
            d = run_custom("run", n, s, dim, nt, flags, "unknown", dt)
            a, m = flat(f"{name}_step{n}_", state_record(d, nt, 7))
            arrays.update(a)
            meta.update(m)
    save("ref_j2_cap", arrays, meta)


if __name__ == "__main__":
    if not os.path.exists(DRIVER):
        oracle.build_ref(verbose=True)
    which = sys.argv[1:] or ["dam_break", "beams", "custom_bggs"]
    for w in which:
        globals()[w]()
