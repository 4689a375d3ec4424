"""Seeded particle sets for the branches of the reference that none of its
shipped cases exercises with defined values, and the configs that state the
reference's per-nick constants for them explicitly.

  - external_force pair branches (src/Physic/externalForce.f90:103-174): the
    fluid-fluid penalty between two fluid types (:126-141) and the
    fluid-solid coupling (:142-170), next to the wall force (:109-124) and
    gravity with its per-nick variants (:27-58);
  - the J2 cap of the deviatoric stress for type 104 in leap_frog
    (src/TimeIntegration/timeIntegration.f90:149-166, :190-205, :225-240).

tests/golden/make_golden.py runs the compiled reference on these sets
(oracle/_ref/ref_driver, case "custom") and commits its outputs; the tests
regenerate the same inputs from the same seeds.
"""
from __future__ import annotations

import numpy as np

from mysph_amd import config as C

# ref_driver custom-state flags
F_VISC, F_GRAV, F_EXF, F_AV, F_KGC, F_SUMD, F_NORM, F_AS = 1, 2, 4, 8, 16, 32, 64, 128

# type 104 (aluminium): eos.f90:240-247 (K = 7.235e10, rho0 = 2700, a
# single-precision literal in a real(8) parameter), viscosity.f90:28
# (G = 2.77443609023e10), yield stress of the J2 cap timeIntegration.f90:161
AL_RHO0 = 2700.0
AL_K = C.f32(7.235e10)
AL_G = C.f32(2.77443609023e10)
AL_SIGMA_Y = C.f32(3e8)


def _lattice(n_per, dx, origin, dim, rng, jitter):
    g = np.stack(np.meshgrid(*[np.arange(k) for k in n_per], indexing="ij"), -1).reshape(-1, dim)
    x = np.asarray(origin, dtype=np.float64) + (g + 0.5) * dx
    return x + jitter * dx * (2 * rng.random(x.shape) - 1)


def _state(x, types, h, c, rho, dx, dim, v=None):
    n = len(x)
    return dict(x=np.ascontiguousarray(x, np.float64),
                v=np.zeros((n, dim)) if v is None else np.ascontiguousarray(v, np.float64),
                mass=np.asarray(rho, np.float64) * dx ** dim * np.ones(n),
                rho=np.asarray(rho, np.float64) * np.ones(n), p=np.zeros(n), e=np.zeros(n),
                c=np.asarray(c, np.float64) * np.ones(n), h=np.asarray(h, np.float64) * np.ones(n),
                type=np.asarray(types, np.int32), state=np.zeros(n, np.int32),
                boundary=np.zeros(n, np.int32))


def _join(parts):
    out = {}
    for k in parts[0]:
        out[k] = np.concatenate([p[k] for p in parts])
    return out


def nick_config(dim: int, nick: str, flags: int, dt: float, x=None) -> C.Config:
    """sphgpu_config with the constants the reference selects for `nick`
    (externalForce.f90:62-101, artificialViscosity.f90:30-79) and the
    switches of `flags`. `x` (particle-major, all particles) gives P(2) for
    the dam_break wall-force range."""
    c = C.base_config(dim)
    c.viscosity = int(bool(flags & F_VISC))
    c.gravity = int(bool(flags & F_GRAV))
    c.external_force = int(bool(flags & F_EXF))
    c.artificial_viscosity = int(bool(flags & F_AV))
    c.kernel_correction = int(bool(flags & F_KGC))
    c.summation_density = int(bool(flags & F_SUMD))
    c.normalize_density = int(bool(flags & F_NORM))
    c.artificial_stress = int(bool(flags & F_AS))
    c.dt = dt
    if nick in ("db_gate", "water_impact"):
        c.gravity_mode = C.GRAVITY_SKIP_SOLID           # externalForce.f90:42-48
    if nick == "db_gate":
        c.couple_delta = C.f32(0.001)                   # :93-95
    elif nick == "water_impact":
        c.couple_delta = C.f32(0.002)                   # :96-97
    elif nick == "dam_break":
        c.lj_factor = 10.0                              # :84-87
        c.lj_r0 = abs(float(x[1, 0]) - float(x[1, 1]))
        c.couple_delta = C.f32(0.2)
    return c


def external_force_sets():
    """Particle sets on which the external force's pair branches fire, with
    defined values for every quantity the routine reads (x, h, SoundSpeed,
    Type): each is (name, nick, dim, flags, ntotal, state dict)."""
    rng = np.random.default_rng(20261206)
    sets = []

    # 2D, default nick: two fluid types (penalty, pe = (h_i + h_j) / 2r >= 1
    # for near pairs), per-particle h, a wall of dummies with two particles
    # closer to it than the default LJ range r0 = 1.25e-5
    dx = 0.01
    xf = _lattice((24, 16), dx, (0.0, 0.03), 2, rng, 0.3)
    t = np.where(xf[:, 0] < 0.12, 2, 3)
    t[rng.random(len(t)) < 0.1] = 3
    h = 1.2 * dx * (1 + 0.2 * rng.random(len(xf)))
    fl = _state(xf, t, h, 30.0 + 5 * rng.random(len(xf)), 1000.0, dx, 2)
    xw = _lattice((24, 3), dx, (0.0, 0.0), 2, rng, 0.0)
    wall = _state(xw, np.full(len(xw), -2), 1.2 * dx, 0.0, 1000.0, dx, 2)
    # two real particles next to wall particles, inside r0
    fl["x"][0] = xw[len(xw) - 24 + 1] + [1.0e-5, 0.4e-5]
    fl["x"][1] = xw[len(xw) - 24 + 7] + [-0.3e-5, 0.9e-5]
    sets.append(("penalty_2d", "unknown", 2, F_GRAV | F_EXF, len(xf), _join([fl, wall])))

    # 2D, nick db_gate (coupling range 0.001): fluid around a solid plate at
    # 0.5 mm spacing, so pairs fall in every branch of the coupling's
    # eta (r / 0.75 / h_mean) and chi (1 - r / delta)
    dx = 0.0005
    xf = _lattice((30, 20), dx, (0.0, 0.0015), 2, rng, 0.4)
    xs = _lattice((6, 20), dx, (0.0060, 0.0015), 2, rng, 0.15)
    keep = (xf[:, 0] < 0.0058) | (xf[:, 0] > 0.0092)
    xf = xf[keep]
    fl = _state(xf, np.full(len(xf), 2), 1.2 * dx, 20.0 + rng.random(len(xf)), 1000.0, dx, 2)
    so = _state(xs, np.full(len(xs), 102), 1.2 * dx * 0.9, 135.0, 1100.0, dx, 2)
    xw = _lattice((30, 3), dx, (0.0, 0.0), 2, rng, 0.0)
    wall = _state(xw, np.full(len(xw), -2), 1.2 * dx, 0.0, 1000.0, dx, 2)
    sets.append(("coupling_2d", "db_gate", 2, F_GRAV | F_EXF,
                 len(xf) + len(xs), _join([fl, so, wall])))

    # 3D, nick water_impact (coupling range 0.002): two fluid types, a type
    # 102 and a type 104 solid block (no coupling between two solids), dummies
    dx = 0.001
    xa = _lattice((10, 6, 8), dx, (0.0, 0.0, 0.002), 3, rng, 0.35)
    ta = np.where(rng.random(len(xa)) < 0.3, 3, 2)
    xb = _lattice((3, 6, 8), dx, (0.010, 0.0, 0.002), 3, rng, 0.2)
    tb = np.where(xb[:, 2] < 0.006, 102, 104)
    fa = _state(xa, ta, 1.2 * dx * (1 + 0.1 * rng.random(len(xa))), 15.0, 1000.0, dx, 3)
    sb = _state(xb, tb, 1.2 * dx, 150.0 + 10 * rng.random(len(xb)), 1100.0, dx, 3)
    xw = _lattice((13, 6, 2), dx, (0.0, 0.0, 0.0), 3, rng, 0.0)
    wall = _state(xw, np.full(len(xw), -2), 1.2 * dx, 0.0, 1000.0, dx, 3)
    sets.append(("mixed_3d", "water_impact", 3, F_GRAV | F_EXF,
                 len(xa) + len(xb), _join([fa, sb, wall])))
    return sets


def j2_sets():
    """Type-104 blocks driven by a velocity field strong enough that the
    J2 cap of the deviatoric stress engages within the run: (name, dim,
    flags, nsteps, dt, state dict). Real particles only (leap_frog counts no
    dummies it does not generate)."""
    rng = np.random.default_rng(20261207)
    out = []
    cs = float(np.sqrt(AL_K / AL_RHO0))
    for dim, n_per, a in ((2, (16, 16), 5000.0), (3, (9, 9, 9), 9000.0)):
        dx = 1e-3
        x = _lattice(n_per, dx, (0.0,) * dim, dim, rng, 0.05)
        xc = x - x.mean(axis=0)
        v = np.zeros_like(x)
        v[:, 0] = a * xc[:, 1]
        v[:, 1] = 0.3 * a * xc[:, 0]
        if dim == 3:
            v[:, 2] = 0.5 * a * xc[:, 0] - 0.2 * a * xc[:, 1]
        s = _state(x, np.full(len(x), 104), 1.2 * dx, cs, AL_RHO0, dx, dim, v=v)
        out.append((f"j2_{dim}d", dim, F_VISC | F_AV, 60, 5e-8, s))
    return out


def j2_config(dim: int, dt: float, flags: int) -> C.Config:
    c = nick_config(dim, "unknown", flags, dt)
    c.add_material(104, C.EOS_LINEAR, AL_RHO0, AL_K / AL_RHO0, 1.0, eta=AL_G, sigma_y=AL_SIGMA_Y)
    return c


def as_case(s: dict, ntotal: int, dim: int, cfg: C.Config) -> dict:
    """A case dict (the layout mysph_amd.cases produces) for a set."""
    n = len(s["h"])
    case = {k: np.array(v, copy=True) for k, v in s.items()}
    case.update(dim=dim, n_real=int(ntotal), n_dummy=int(n - ntotal), config=cfg)
    return case


def write_driver_input(path, s: dict, dim: int, skf: int, ntotal: int, kpair: int, flags: int,
                       nick: str, dt: float, pa_sph: int = 2, sle: int = 0):
    """The custom-state input of ref_driver (oracle/ref_driver.f90,
    read_custom_state)."""
    n = len(s["h"])
    with open(path, "wb") as f:
        np.array([dim, skf, ntotal, n, kpair, flags, pa_sph, sle], "<i4").tofile(f)
        f.write(nick.ljust(16).encode())
        np.array([dt], "<f8").tofile(f)
        for k in ("x", "v"):
            np.ascontiguousarray(s[k], "<f8").ravel().tofile(f)
        for k in ("mass", "rho", "p", "e", "c", "h"):
            np.ascontiguousarray(s[k], "<f8").tofile(f)
        for k in ("type", "state", "boundary"):
            np.ascontiguousarray(s[k], "<i4").tofile(f)
