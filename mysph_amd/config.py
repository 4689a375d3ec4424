"""Run configuration: the ctypes mirror of ``sphgpu_config`` (include/sphgpu.h)
and presets that reproduce the reference's per-case constants.

The reference selects most physics constants by the case ``nick`` string
(src/Physic/eos.f90:35-40, src/Numeric/artificialViscosity.f90:30-79,
src/Physic/externalForce.f90:62-101, src/Physic/viscosity.f90:14-33). Here each
constant is an explicit field; ``reference_*`` presets set them to the exact
values the reference uses for that nick, including single-precision literals
that the reference promotes to double (``f32``).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

MAX_MATERIALS = 8

# sphgpu_precision
FP64 = 0
MIXED = 1

# sphgpu_eos
EOS_NONE, EOS_TAIT, EOS_LINEAR, EOS_ISOTHERMAL = 0, 1, 2, 3

# sphgpu_gravity_mode
GRAVITY_ALL, GRAVITY_SKIP_SOLID = 0, 1

# sphgpu_field
F_X, F_V, F_RHO, F_P, F_E, F_MASS, F_H, F_C, F_DISP, F_STRESS, F_SHEAR, \
    F_DVDT, F_DRHODT, F_DEDT, F_DSDT, F_DIVX, F_DIVV, F_ETA = range(18)

FIELD_NAMES = {
    "x": F_X, "v": F_V, "rho": F_RHO, "p": F_P, "e": F_E, "mass": F_MASS,
    "h": F_H, "c": F_C, "disp": F_DISP, "stress": F_STRESS, "shear": F_SHEAR,
    "dvdt": F_DVDT, "drhodt": F_DRHODT, "dedt": F_DEDT, "dsdt": F_DSDT,
    "divx": F_DIVX, "divv": F_DIVV, "eta": F_ETA,
}


def field_components(field: int, dim: int) -> int:
    if field in (F_X, F_V, F_DISP, F_DVDT):
        return dim
    if field in (F_STRESS, F_SHEAR, F_DSDT):
        return dim * dim
    return 1


def f32(x: float) -> float:
    """Value of a Fortran default-real literal promoted to real(8)."""
    return float(np.float32(x))


class Material(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int32),
        ("eos", ctypes.c_int32),
        ("rho0", ctypes.c_double),
        ("K", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("eta", ctypes.c_double),
        ("sigma_y", ctypes.c_double),
    ]


class Config(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("skf", ctypes.c_int32),
        ("pa_sph", ctypes.c_int32),
        ("summation_density", ctypes.c_int32),
        ("normalize_density", ctypes.c_int32),
        ("viscosity", ctypes.c_int32),
        ("gravity", ctypes.c_int32),
        ("gravity_mode", ctypes.c_int32),
        ("external_force", ctypes.c_int32),
        ("artificial_viscosity", ctypes.c_int32),
        ("artificial_stress", ctypes.c_int32),
        ("kernel_correction", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("n_materials", ctypes.c_int32),
        ("gravity_g", ctypes.c_double),
        ("av_alpha", ctypes.c_double),
        ("av_beta", ctypes.c_double),
        ("av_psi", ctypes.c_double),
        ("lj_factor", ctypes.c_double),
        ("lj_r0", ctypes.c_double),
        ("lj_p1", ctypes.c_double),
        ("lj_p2", ctypes.c_double),
        ("pen_factor", ctypes.c_double),
        ("pen_n1", ctypes.c_double),
        ("pen_n2", ctypes.c_double),
        ("couple_delta", ctypes.c_double),
        ("as_lambda", ctypes.c_double),
        ("as_delta", ctypes.c_double),
        ("as_exponent", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("mat", Material * MAX_MATERIALS),
        ("fluid_visc_skip_solid", ctypes.c_int32),
        ("sle", ctypes.c_int32),
    ]

    def add_material(self, type_: int, eos: int, rho0: float, K: float,
                     gamma: float = 1.0, eta: float = 0.0, sigma_y: float = 0.0):
        k = self.n_materials
        if k >= MAX_MATERIALS:
            raise ValueError("too many materials")
        m = self.mat[k]
        m.type, m.eos, m.rho0, m.K, m.gamma, m.eta, m.sigma_y = (
            abs(int(type_)), int(eos), float(rho0), float(K), float(gamma),
            float(eta), float(sigma_y))
        self.n_materials = k + 1
        return self

    def copy(self) -> "Config":
        c = Config()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self), ctypes.sizeof(Config))
        return c

    def material(self, type_: int):
        a = abs(int(type_))
        for k in range(self.n_materials):
            if self.mat[k].type == a:
                return self.mat[k]
        return None


def base_config(dim: int) -> Config:
    """The reference's defaults for a nick that matches no special case
    (getValueFromToml.f90:99-150, externalForce.f90:62-72,
    artificialViscosity.f90:282-285)."""
    c = Config()
    c.dim = dim
    c.skf = 1
    c.pa_sph = 2
    c.summation_density = 1
    c.gravity_g = -9.81
    c.av_alpha, c.av_beta, c.av_psi = 1.0, 1.0, f32(0.1)
    c.lj_factor, c.lj_r0, c.lj_p1, c.lj_p2 = f32(1e-2), f32(1.25e-5), 12.0, 4.0
    c.pen_factor, c.pen_n1, c.pen_n2 = 1e5, 6.0, 4.0
    c.couple_delta = f32(0.001)
    c.as_lambda, c.as_delta, c.as_exponent = f32(0.2), 0.0, 4.0
    c.dt = 5e-4
    c.precision = FP64
    return c


def tait_B(c0: float, rho0: float = 1000.0, gamma: float = 7.0) -> float:
    """B = c^2 rho0 / gamma as eos.f90:41 evaluates it (c**2 * rho0 / gamma)."""
    return c0 * c0 * rho0 / gamma


def reference_dam_break(r0: float) -> Config:
    """nick "dam_break" with data/damBreak/sph.toml. ``r0`` is the wall-force
    range the reference takes from particle 2 on first entry
    (externalForce.f90:84-87: abs(P(2)%x(1) - P(2)%x(2)))."""
    c = base_config(2)
    c.summation_density = 0
    c.viscosity = 1
    c.gravity = 1
    c.external_force = 1
    c.artificial_viscosity = 1
    c.lj_factor, c.lj_r0 = 10.0, float(r0)
    c.couple_delta = f32(0.2)
    c.dt = 1e-3
    # type 2 water: Tait with c = 50 for dam_break (eos.f90:37), eta = 1e-3 (viscosity.f90:18)
    c.add_material(2, EOS_TAIT, 1000.0, tait_B(50.0), 7.0, eta=f32(1e-3))
    return c


def reference_can_beam(dim: int = 2) -> Config:
    """nick "can_beam" / "can_beam_3d" with data/cantileverBeam*/sph.toml:
    solid type 102 (eos.f90:219-226, viscosity.f90:24), AV 0.5/0.5/0.1
    (artificialViscosity.f90:47-50), kernel gradient correction on."""
    c = base_config(dim)
    c.summation_density = 0
    c.viscosity = 1
    c.gravity = 1
    c.external_force = 1
    c.artificial_viscosity = 1
    c.kernel_correction = 1
    c.av_alpha, c.av_beta, c.av_psi = 0.5, 0.5, f32(0.1)
    c.as_delta = f32(1e-3) if dim == 2 else 0.0   # artificialTerm.f90:30-31
    c.dt = 2e-6
    c.add_material(102, EOS_LINEAR, 1100.0, f32(2e7) / 1100.0, 1.0, eta=f32(4.27e6))
    return c


def dam_break_config(dim: int, column_height: float, h: float, alpha: float = 0.1,
                     beta: float = 0.0, psi: float = 0.1, cfl: float = 0.2,
                     precision: int = FP64) -> Config:
    """The 2D/3D dam-break benchmark physics: Tait water with
    c0 = 10 sqrt(2 g H), Monaghan AV, gravity, continuity density, fixed
    boundary particles, fixed dt = cfl * h / (c0 (1 + 1.2 alpha)) (the
    Courant form of courantNumber.f90:27-29 at rest)."""
    c = base_config(dim)
    c.summation_density = 0
    c.viscosity = 0
    c.gravity = 1
    c.external_force = 1
    c.artificial_viscosity = 1
    c.av_alpha, c.av_beta, c.av_psi = alpha, beta, psi
    c.lj_r0 = 0.0  # no wall repulsion: the fixed particles carry the wall
    c0 = 10.0 * math.sqrt(2.0 * 9.81 * column_height)
    c.dt = cfl * h / (c0 * (1.0 + 1.2 * alpha))
    c.precision = precision
    c.add_material(2, EOS_TAIT, 1000.0, tait_B(c0), 7.0, eta=f32(1e-3))
    return c


def gate_config(dim: int, column_height: float, h_min: float, E: float = 12e6, nu: float = 0.4,
                rho_s: float = 1100.0, couple_delta: float = 0.001, alpha: float = 0.1,
                beta: float = 0.0, psi: float = 0.1, cfl: float = 0.2,
                precision: int = FP64) -> Config:
    """Dam break with an elastic gate: the benchmark water (Tait,
    c0 = 10 sqrt(2 g H)) and a Hookean solid of type 102 with K = E / (3 (1 -
    2 nu)) and G = E / (2 (1 + nu)) (E = 12 MPa, nu = 0.4 give the
    reference's type 102, eos.f90:219-226 and viscosity.f90:24), the
    fluid-solid coupling force within ``couple_delta`` (externalForce.f90:
    142-170; 0.001 for nick "db_gate", :93-95), viscosity on, which the
    solid's strain-rate pass needs (internalForce.f90:48-100;
    data/damBreakwithElasticGate/sph.toml), and fixed dt =
    cfl * h_min / (c_max (1 + 1.2 alpha)) over both sound speeds.

    fluid_visc_skip_solid is on: with the reference's pair term a fluid
    next to the gate takes eta_j = G (4.3 MPa) times the solid's strain
    rate as a viscous stress, and the run diverges within a few steps at
    any practical dt."""
    c = base_config(dim)
    c.summation_density = 0
    c.viscosity = 1
    c.gravity = 1
    c.external_force = 1
    c.artificial_viscosity = 1
    c.av_alpha, c.av_beta, c.av_psi = alpha, beta, psi
    c.lj_r0 = 0.0
    c.couple_delta = couple_delta
    c0 = 10.0 * math.sqrt(2.0 * 9.81 * column_height)
    K = E / (3.0 * (1.0 - 2.0 * nu))
    G = E / (2.0 * (1.0 + nu))
    cs = math.sqrt(K / rho_s)
    c.dt = cfl * h_min / (max(c0, cs) * (1.0 + 1.2 * alpha))
    c.precision = precision
    c.add_material(2, EOS_TAIT, 1000.0, tait_B(c0), 7.0, eta=f32(1e-3))
    c.add_material(102, EOS_LINEAR, rho_s, K / rho_s, 1.0, eta=G)
    c.fluid_visc_skip_solid = 1
    return c


def sound_speed(cfg: Config, type_: int = 2) -> float:
    m = cfg.material(type_)
    if m is None:
        return 0.0
    if m.eos == EOS_TAIT:
        return math.sqrt(m.K * m.gamma / m.rho0)
    if m.eos == EOS_LINEAR:
        return math.sqrt(m.K)
    return 0.0
