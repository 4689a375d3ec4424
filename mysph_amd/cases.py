"""Case set-up: initial particle sets on the host.

Reference cases are reproduced bit for bit (same lattice arithmetic, same
single-precision literals, same inverse-EOS call):
  ``reference_dam_break``  src/Initialize/input.f90:1031-1088 (fluid) and
                           src/Physic/dummyParticles.f90:544-659 (4-layer walls,
                           baffle, lid)
  ``reference_can_beam``   src/Initialize/input.f90:1336-1366
  ``reference_can_beam_3d`` src/Initialize/input.f90:1368-1405
The benchmark dam breaks (2D and 3D, 3-layer fixed boundary, seeded jitter)
and the elastic beam are new cases on the same conventions.

A case is a dict of numpy arrays over n = n_real + n_dummy particles
(particle-major vectors) plus ``n_real``, ``n_dummy``, ``dim`` and ``config``.
"""
from __future__ import annotations

import math

import numpy as np

from . import config as C

f32 = C.f32
SEED_JITTER = 20261203

_M64 = (1 << 64) - 1


def splitmix64(keys: np.ndarray) -> np.ndarray:
    """SplitMix64 output function of (key + golden gamma), vectorised over
    uint64 keys."""
    z = (keys.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def jitter_unit(ids: np.ndarray, dim: int, seed: int = SEED_JITTER) -> np.ndarray:
    """Counter-based uniform values in [-1, 1) for (particle id, axis):
    key = seed XOR (4*id + axis), value = 2 * (top 53 bits / 2^53) - 1."""
    ids = np.asarray(ids, dtype=np.uint64)
    d = np.arange(dim, dtype=np.uint64)
    keys = np.uint64(seed) ^ ((ids[:, None] << np.uint64(2)) | d[None, :])
    z = splitmix64(keys.ravel())
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return (2.0 * u - 1.0).reshape(len(ids), dim)


def _tait_inverse(p, B, rho0=1000.0, gamma=7.0):
    """rho = (p/B + 1)**(1/gamma) * rho0, eos.f90:46-47, with the C library's
    pow per value: numpy's vectorised power may take a SIMD path whose last
    bit differs from libm's and from one host CPU to another, while libm pow
    is what the Fortran host and the reference call."""
    e = 1.0 / gamma
    base = np.asarray(p, dtype=np.float64) / B + 1.0
    return np.fromiter((math.pow(b, e) for b in base), dtype=np.float64, count=base.size) * rho0


def _empty(n: int, dim: int) -> dict:
    return {
        "x": np.zeros((n, dim)), "v": np.zeros((n, dim)), "mass": np.zeros(n),
        "rho": np.zeros(n), "p": np.zeros(n), "e": np.zeros(n), "c": np.zeros(n),
        "h": np.zeros(n), "type": np.zeros(n, dtype=np.int32),
        "state": np.zeros(n, dtype=np.int32), "boundary": np.zeros(n, dtype=np.int32),
    }


def _concat(parts: list, dim: int) -> dict:
    out = {}
    for k in parts[0]:
        out[k] = np.concatenate([p[k] for p in parts], axis=0)
    return out


# ---------------------------------------------------------------- reference

def reference_dam_break() -> dict:
    """The reference 2D dam break: 124 x 124 fluid (dx = 0.2f, so
    int(25/dx) = 124) plus 5734 regenerated-each-step wall particles, which
    are static and therefore generated once."""
    dx = f32(0.2)
    g981 = f32(9.81)
    B = C.tait_B(50.0)
    nx = int((25.0 - 0.0) / dx)
    ny = int((25.0 - 0.0) / dx)
    i = np.repeat(np.arange(1, nx + 1), ny)
    j = np.tile(np.arange(1, ny + 1), nx)
    fl = _empty(nx * ny, 2)
    # (i-0.5) is a default-real expression: exact in float32 for these i
    fl["x"][:, 0] = 0.0 + (i.astype(np.float32) - np.float32(0.5)).astype(np.float64) * dx
    fl["x"][:, 1] = 0.0 + (j.astype(np.float32) - np.float32(0.5)).astype(np.float64) * dx
    fl["rho"][:] = 1000.0
    fl["mass"][:] = 1000.0 * dx * dx
    fl["p"][:] = 1000.0 * g981 * (25.0 - fl["x"][:, 1])
    fl["c"][:] = 50.0
    fl["type"][:] = 2
    fl["h"][:] = dx
    fl["rho"][:] = _tait_inverse(fl["p"], B)

    # dummyParticles.f90:544-659; dx re-derived from particle 2 (:561-564)
    ddx = abs(fl["x"][1, 0] - fl["x"][1, 1])
    wall = (0.0, 100.0, 0.0, 40.0)
    layer = 4
    xs, rho, mass, p, hydro = [], [], [], [], []

    def add(x, y, m, hyd):
        xs.append((x, y)); mass.append(m); hydro.append(hyd)

    for l in range(1, layer + 1):
        Ny = math.floor((wall[3] - wall[2]) / ddx) + layer
        lf = float(np.float32(l) - np.float32(0.5))
        for ii in range(1, Ny + 1):
            yy = wall[2] + float(np.float32(ii - layer) - np.float32(0.5)) * ddx
            add(wall[0] - lf * ddx, yy, 1000.0 * ddx * ddx, True)
        for ii in range(1, Ny + 1):
            yy = wall[2] + float(np.float32(ii - layer) - np.float32(0.5)) * ddx
            add(wall[1] + lf * ddx, yy, 1000.0 * ddx * ddx, True)
        Nx = math.floor((wall[1] - wall[0]) / ddx) + 1
        for ii in range(1, Nx + 1):
            xx = wall[0] + float(np.float32(ii) - np.float32(0.5)) * ddx
            add(xx, wall[2] - lf * ddx, 1000.0 * ddx * ddx, None)
    Nx = math.floor(3 / (ddx / 2))
    for l in range(1, layer + 1):
        for ii in range(1, Nx - l + 2):
            xx = (wall[0] + wall[1]) / 2 + float(np.float32(ii + l) - np.float32(1.5)) * ddx / 2
            yy = wall[2] + float(np.float32(ii) - np.float32(0.5)) * ddx / 2
            add(xx, yy, 1000.0 * ddx * ddx / 4, True)
    Nx = math.floor((wall[1] - wall[0]) / ddx) + 1
    for l in range(1, layer + 1):
        lf = float(np.float32(l) + np.float32(0.5))
        for ii in range(1, Nx + 1):
            xx = wall[0] + float(np.float32(ii) - np.float32(0.5)) * ddx
            add(xx, wall[3] - lf * ddx, 1000.0 * ddx * ddx, False)
    nd = len(xs)
    du = _empty(nd, 2)
    du["x"][:] = np.array(xs)
    du["mass"][:] = np.array(mass)
    du["rho"][:] = 1000.0
    du["type"][:] = -2
    du["h"][:] = ddx
    for k in range(nd):
        if hydro[k] is True:
            pk = 1000.0 * g981 * max(25.0 - du["x"][k, 1], 0.0)
            du["p"][k] = pk
            du["rho"][k] = math.pow(pk / B + 1.0, 1.0 / 7.0) * 1000.0
    case = _concat([fl, du], 2)
    case.update(n_real=nx * ny, n_dummy=nd, dim=2, config=C.reference_dam_break(r0=ddx),
                name="reference_dam_break")
    return case


def reference_can_beam() -> dict:
    """data/cantileverBeam: 101 x 5 solid (type 102), clamped first column."""
    dx = f32(1e-3)
    bulk, rho0 = f32(2e7), 1100.0
    c = math.sqrt(bulk / rho0)
    nx, ny = 101, 5
    i = np.repeat(np.arange(1, nx + 1), ny)
    j = np.tile(np.arange(1, ny + 1), nx)
    s = _empty(nx * ny, 2)
    s["x"][:, 0] = (i - 1) * dx
    s["x"][:, 1] = (j - 1 - ny // 2) * dx
    s["rho"][:] = rho0
    s["mass"][:] = rho0 * (dx * dx)
    s["c"][:] = c
    s["type"][:] = 102
    s["h"][:] = dx * 2
    s["boundary"][:] = (i == 1).astype(np.int32)
    s.update(n_real=nx * ny, n_dummy=0, dim=2, config=C.reference_can_beam(2),
             name="reference_can_beam")
    return s


def reference_can_beam_3d() -> dict:
    """data/cantileverBeam3D: 201 x 5 x 9 solid, clamped first layer."""
    dx = f32(5e-4)
    bulk, rho0 = f32(2e7), 1100.0
    c = math.sqrt(bulk / rho0)
    nx, ny, nz = 201, 5, 9
    ii, jj, kk = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), np.arange(1, nz + 1),
                             indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    s = _empty(nx * ny * nz, 3)
    s["x"][:, 0] = (ii - 1) * dx
    s["x"][:, 1] = (jj - 1 - ny // 2) * dx
    s["x"][:, 2] = (kk - 1 - nz // 2) * dx
    s["rho"][:] = rho0
    s["mass"][:] = rho0 * (dx * dx * dx)
    s["c"][:] = c
    s["type"][:] = 102
    s["h"][:] = dx * 2
    s["boundary"][:] = (ii == 1).astype(np.int32)
    s.update(n_real=len(ii), n_dummy=0, dim=3, config=C.reference_can_beam(3),
             name="reference_can_beam_3d")
    return s


# ----------------------------------------------------------- benchmark cases

def dam_break(dim: int, dx: float, column=(1.0, 2.0), tank=(4.0, 3.0), layers: int = 3,
              h_over_dx: float = 1.2, jitter: float = 0.01, precision: int = C.FP64,
              alpha: float = 0.1, seed: int = SEED_JITTER) -> dict:
    """Dam break with a fixed ``layers``-deep boundary lattice at dx, offset
    half a spacing outside the tank, and no lid.

    2D: ``column`` = (width, height), ``tank`` = (length, wall height).
    3D: ``column`` = (x, y, z), ``tank`` = (length, width, wall height).
    Fluid lattice ((i+1/2) dx per axis) plus uniform jitter of +-jitter*dx
    from ``jitter_unit``. Hydrostatic initial pressure and inverse-Tait
    density for fluid and boundary, as the reference sets its water and
    walls (input.f90:1053-1058, dummyParticles.f90:578-584)."""
    H = column[-1]
    h = h_over_dx * dx
    cfg = C.dam_break_config(dim, H, h, alpha=alpha, precision=precision)
    mat = cfg.material(2)
    B = mat.K
    c0 = C.sound_speed(cfg)
    g = 9.81
    ncol = [int(round(L / dx)) for L in column]
    ntank = [int(round(L / dx)) for L in tank]
    # fluid
    grids = np.meshgrid(*[np.arange(n) for n in ncol], indexing="ij")
    idx = np.stack([gr.ravel() for gr in grids], axis=1)
    nf = idx.shape[0]
    xf = (idx + 0.5) * dx
    if jitter:
        xf = xf + jitter * dx * jitter_unit(np.arange(nf), dim, seed)
    fl = _empty(nf, dim)
    fl["x"][:] = xf
    fl["mass"][:] = 1000.0 * dx ** dim
    fl["p"][:] = 1000.0 * g * np.maximum(H - xf[:, -1], 0.0)
    fl["rho"][:] = _tait_inverse(fl["p"], B)
    fl["c"][:] = c0
    fl["h"][:] = h
    fl["type"][:] = 2
    # boundary: lattice cells outside the open-top tank interior, within
    # `layers` of it, up to the wall height
    lo = [-layers] * dim
    hi = [ntank[d] + layers for d in range(dim - 1)] + [ntank[-1]]
    grids = np.meshgrid(*[np.arange(lo[d], hi[d]) for d in range(dim)], indexing="ij")
    bidx = np.stack([gr.ravel() for gr in grids], axis=1)
    inside = np.ones(len(bidx), dtype=bool)
    for d in range(dim - 1):
        inside &= (bidx[:, d] >= 0) & (bidx[:, d] < ntank[d])
    inside &= bidx[:, -1] >= 0
    bidx = bidx[~inside]
    nb = len(bidx)
    xb = (bidx + 0.5) * dx
    bd = _empty(nb, dim)
    bd["x"][:] = xb
    bd["mass"][:] = 1000.0 * dx ** dim
    bd["p"][:] = 1000.0 * g * np.maximum(H - xb[:, -1], 0.0)
    bd["rho"][:] = _tait_inverse(bd["p"], B)
    bd["c"][:] = c0
    bd["h"][:] = h
    bd["type"][:] = -2
    case = _concat([fl, bd], dim)
    case.update(n_real=nf, n_dummy=nb, dim=dim, config=cfg,
                name=f"dam_break_{dim}d_dx{dx:g}")
    return case


def dam_break_2d(dx: float = 0.02, **kw) -> dict:
    """2D dam break: 1.0 x 2.0 m column in a 4.0 x 3.0 m tank."""
    return dam_break(2, dx, column=(1.0, 2.0), tank=(4.0, 3.0), **kw)


def dam_break_3d(dx: float = 0.0075, **kw) -> dict:
    """3D dam break: 0.6 x 0.6 x 1.2 m column in a 1.6 x 0.6 m tank with
    walls to 1.3 m (the column is taller than a 1.0 m box, so the walls are
    raised), no lid. dx = 0.0075 gives 80 x 80 x 160 = 1,024,000 fluid."""
    return dam_break(3, dx, column=(0.6, 0.6, 1.2), tank=(1.6, 0.6, 1.3), **kw)


def dam_break_3d_fluid_count(dx: float) -> int:
    return int(round(0.6 / dx)) ** 2 * int(round(1.2 / dx))


def dam_break_gate_3d(dx_f: float = 0.0016, dx_s: float = 0.002, width: float = 0.6,
                      column=(0.1, 0.14), tank=(0.2, 0.15), gate=(0.005, 0.1), layers: int = 3,
                      h_over_dx: float = 1.2, jitter: float = 0.01, precision: int = C.FP64,
                      alpha: float = 0.1, seed: int = SEED_JITTER) -> dict:
    """3D dam break with an elastic gate: the reference's 2D "db_gate"
    layout (input.f90:1090-1154, dummyParticles.f90:661-765) extruded over
    ``width`` in y, at two resolutions.

    Water column x in [0, column[0]], z in [0, column[1]] on a dx_f lattice
    (jittered as the other benchmark cases). Gate: a type-102 plate of
    int(gate[0] / dx_s) + 1 layers at x = column[0] + (i + 1/2) dx_s, rows
    z = (j + 3/2) dx_s for j < int(gate[1] / dx_s) + 1 (a gap above the floor,
    as input.f90:1138 leaves one), the top row clamped (Boundary = 1, as
    input.f90:1147-1149). Fixed boundary (type -2, ``layers`` deep, dx_f):
    floor, the four tank walls to tank[1], and the wall above the gate from
    one dx_f over its clamped row to tank[1] (the reference's "Middle Wall").
    h = h_over_dx * dx per phase: the smoothing length varies by particle.
    The defaults give 2,046,000 fluid and 45,900 solid particles."""
    dim = 3
    L, Hw = tank
    Wc, Hc = column
    h_f, h_s = h_over_dx * dx_f, h_over_dx * dx_s
    cfg = C.gate_config(dim, Hc, min(h_f, h_s), alpha=alpha, couple_delta=dx_f, precision=precision)
    B = cfg.material(2).K
    c0 = C.sound_speed(cfg, 2)
    cs = C.sound_speed(cfg, 102)
    g = 9.81
    # water
    n = [int(round(Wc / dx_f)), int(round(width / dx_f)), int(round(Hc / dx_f))]
    grids = np.meshgrid(*[np.arange(k) for k in n], indexing="ij")
    idx = np.stack([gr.ravel() for gr in grids], axis=1)
    nf = idx.shape[0]
    xf = (idx + 0.5) * dx_f
    if jitter:
        xf = xf + jitter * dx_f * jitter_unit(np.arange(nf), dim, seed)
    fl = _empty(nf, dim)
    fl["x"][:] = xf
    fl["mass"][:] = 1000.0 * dx_f ** 3
    fl["p"][:] = 1000.0 * g * np.maximum(Hc - xf[:, 2], 0.0)
    fl["rho"][:] = _tait_inverse(fl["p"], B)
    fl["c"][:] = c0
    fl["h"][:] = h_f
    fl["type"][:] = 2
    # gate
    ns = [int(gate[0] / dx_s) + 1, int(round(width / dx_s)), int(gate[1] / dx_s) + 1]
    grids = np.meshgrid(*[np.arange(k) for k in ns], indexing="ij")
    sidx = np.stack([gr.ravel() for gr in grids], axis=1)
    xs = np.empty((len(sidx), 3))
    xs[:, 0] = Wc + (sidx[:, 0] + 0.5) * dx_s
    xs[:, 1] = (sidx[:, 1] + 0.5) * dx_s
    xs[:, 2] = (sidx[:, 2] + 1.5) * dx_s
    so = _empty(len(sidx), dim)
    so["x"][:] = xs
    so["rho"][:] = 1100.0
    so["mass"][:] = 1100.0 * dx_s ** 3
    so["c"][:] = cs
    so["h"][:] = h_s
    so["type"][:] = 102
    so["boundary"][:] = (sidx[:, 2] == ns[2] - 1).astype(np.int32)
    gate_x1 = xs[:, 0].max() + 0.5 * dx_s
    gate_z1 = xs[:, 2].max()
    # fixed boundary on the dx_f lattice
    nt = [int(round(L / dx_f)), int(round(width / dx_f)), int(round(Hw / dx_f))]
    lo = [-layers, -layers, -layers]
    hi = [nt[0] + layers, nt[1] + layers, nt[2]]
    grids = np.meshgrid(*[np.arange(lo[d], hi[d]) for d in range(dim)], indexing="ij")
    bidx = np.stack([gr.ravel() for gr in grids], axis=1)
    xb = (bidx + 0.5) * dx_f
    inside = (bidx[:, 0] >= 0) & (bidx[:, 0] < nt[0]) & (bidx[:, 1] >= 0) & (bidx[:, 1] < nt[1]) & \
             (bidx[:, 2] >= 0)
    middle = inside & (xb[:, 0] > Wc) & (xb[:, 0] < gate_x1) & (xb[:, 2] > gate_z1 + dx_f)
    keep = ~inside | middle
    xb = xb[keep]
    bd = _empty(len(xb), dim)
    bd["x"][:] = xb
    bd["mass"][:] = 1000.0 * dx_f ** 3
    bd["p"][:] = 1000.0 * g * np.maximum(Hc - xb[:, 2], 0.0)
    bd["rho"][:] = _tait_inverse(bd["p"], B)
    bd["c"][:] = c0
    bd["h"][:] = h_f
    bd["type"][:] = -2
    case = _concat([fl, so, bd], dim)
    case.update(n_real=nf + len(xs), n_dummy=len(xb), dim=dim, config=cfg,
                name=f"dam_break_gate_3d_dx{dx_f:g}_{dx_s:g}", n_fluid=nf, n_solid=len(xs))
    return case


def elastic_beam(dx: float = 0.0005, length: float = 0.2, thickness: float = 0.02,
                 E: float = 2e9, nu: float = 0.3975, rho0: float = 1000.0,
                 amplitude: float = 0.01, clamp_columns: int = 3,
                 h_over_dx: float = 1.2, precision: int = C.FP64) -> dict:
    """2D elastic cantilever: (length x thickness) solid lattice plus
    ``clamp_columns`` fixed columns at x < 0 (Boundary = 1, integrated with
    zero acceleration as fixedBoundary does, boundaryCondition.f90:32-45).
    Initial transverse velocity of the first bending mode of a clamped-free
    beam, amplitude ``amplitude`` at the tip. Hookean solid: K = E/(3(1-2nu)),
    G = E/(2(1+nu)), artificial stress eps = 0.3 as (R_i + R_j) f^4."""
    K = E / (3.0 * (1.0 - 2.0 * nu))
    G = E / (2.0 * (1.0 + nu))
    c = math.sqrt(K / rho0)
    cfg = C.base_config(2)
    cfg.summation_density = 0
    cfg.viscosity = 1
    cfg.gravity = 0
    cfg.external_force = 1
    cfg.artificial_viscosity = 1
    cfg.artificial_stress = 1
    cfg.kernel_correction = 0
    cfg.av_alpha, cfg.av_beta, cfg.av_psi = 1.0, 1.0, 0.1
    cfg.as_lambda, cfg.as_delta, cfg.as_exponent = 0.3, dx, 4.0
    h = h_over_dx * dx
    cfg.dt = 0.2 * h / (c * 1.2)
    cfg.precision = precision
    cfg.add_material(102, C.EOS_LINEAR, rho0, K / rho0, 1.0, eta=G)
    nx = int(round(length / dx))
    ny = int(round(thickness / dx))
    ii, jj = np.meshgrid(np.arange(-clamp_columns, nx), np.arange(ny), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    n = len(ii)
    s = _empty(n, 2)
    s["x"][:, 0] = (ii + 0.5) * dx
    s["x"][:, 1] = (jj + 0.5) * dx - thickness / 2
    # first clamped-free bending mode, kL = 1.875104
    kL = 1.8751040687119611
    sig = (math.cosh(kL) + math.cos(kL)) / (math.sinh(kL) + math.sin(kL))
    xi = np.clip(s["x"][:, 0], 0.0, None) * kL / length

    def mode(z):
        return np.cosh(z) - np.cos(z) - sig * (np.sinh(z) - np.sin(z))

    s["v"][:, 1] = amplitude * mode(xi) / mode(np.array([kL]))[0]
    s["rho"][:] = rho0
    s["mass"][:] = rho0 * dx * dx
    s["c"][:] = c
    s["h"][:] = h
    s["type"][:] = 102
    s["boundary"][:] = (ii < 0).astype(np.int32)
    s.update(n_real=n, n_dummy=0, dim=2, config=cfg, name="elastic_beam")
    return s
