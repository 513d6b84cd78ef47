"""Synthetic workloads: mesh, numerics, boundary conditions, freestream.

ramp  -- inviscid supersonic compression ramp, 48x16x4 hexahedra on
         [0,3]x[0,1]x[0,0.25], bottom sheared 15 deg for x > 0.5, Mach 2.5.
plate -- laminar flat plate, 256x128x64 hexahedra, wall-normal growth 1.08,
         +-10 % seeded interior jitter, reverse Cuthill-McKee order, Mach 0.3,
         Re_L = 1e5, adiabatic no-slip plate for x > 0.
bump  -- tetrahedral (six Kuhn tetrahedra per hexahedron) channel with a
         Gaussian bottom bump, +-5 % jitter, Mach 0.8, Re = 1e6.
ramp_sa, plate_sa, bump_sa -- the same geometries with viscous flow and the
         Spalart-Allmaras module (ramp: no-slip bottom, Re_L = 5e6; plate at
         Re_L = 5e6); turbulence(case) gives the module's parameter block.

Every case accepts a size override so the same physics runs at parity-test
sizes (a few thousand cells) and at benchmark sizes.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

from . import (BC, GRADIENT, INFLOW_VELOCITY, INVISCID_FLOW, LIMITER, MUSCL, PRECOND_SGS, SGS_MULTICOLOUR,
               VISCOUS_FLOW, Bc, LinParams, Numerics)

GAMMA = 1.4
R_AIR = 287.052874  # include/cfd/core/types.hpp:39
PRANDTL = 0.71


def sutherland(T):
    return 1.716e-5 * (T / 273.15) ** 1.5 * (273.15 + 110.4) / (T + 110.4)


def sound_speed(T):
    return math.sqrt(GAMMA * R_AIR * T)


@dataclass
class Case:
    name: str
    mesh: dict
    numerics: Numerics
    bcs: list
    rho: float
    velocity: tuple
    p: float
    linparams: LinParams
    info: dict = field(default_factory=dict)


def make_bc(kind, **kw):
    b = Bc()
    b.type = BC[kind]
    b.inflow_mode = kw.pop("inflow_mode", INFLOW_VELOCITY)
    b.p = kw.pop("p", 101325.0)
    b.t = kw.pop("t", 288.15)
    b.tmp_grad = kw.pop("tmp_grad", 0.0)
    vel = kw.pop("velocity", (0.0, 0.0, 0.0))
    for i in range(3):
        b.velocity[i] = vel[i]
    b.mach = kw.pop("mach", 0.0)
    b.alpha_deg = kw.pop("alpha_deg", 0.0)
    b.beta_deg = kw.pop("beta_deg", 0.0)
    d = kw.pop("direction", (1.0, 0.0, 0.0))
    for i in range(3):
        b.direction[i] = d[i]
    if kw:
        raise ValueError(f"unknown bc keys {sorted(kw)}")
    return b


def make_numerics(viscous, gradient="LEAST_SQUARES_FACE", limiter="VENKAT", venkat_k=0.5, cfl=10.0,
                  reconstruction=MUSCL):
    n = Numerics()
    n.flow_model = VISCOUS_FLOW if viscous else INVISCID_FLOW
    n.gamma = GAMMA
    n.gas_constant = R_AIR
    n.prandtl = PRANDTL
    n.prandtl_turb = 0.9
    n.reconstruction = reconstruction
    n.limiter = LIMITER[limiter]
    n.venkat_k = venkat_k
    n.gradient = GRADIENT[gradient]
    n.cfl = cfl
    return n


def default_linparams(rel_tol=0.1, abs_tol=1e-15, max_iter=100, order=SGS_MULTICOLOUR):
    lp = LinParams()
    lp.preconditioner = PRECOND_SGS
    lp.sgs_order = order
    lp.rel_tol = rel_tol
    lp.abs_tol = abs_tol
    lp.max_iter = max_iter
    return lp


def ramp(ni=48, nj=16, nk=4, mach=2.5, T=300.0, p=101325.0, **num_kw):
    a = sound_speed(T)
    u = mach * a
    rho = p / (R_AIR * T)
    mesh = dict(kind=0, ni=ni, nj=nj, nk=nk, x0=0.0, x1=3.0, y0=0.0, y1=1.0, z0=0.0, z1=0.25,
                ramp_x=0.5, ramp_deg=15.0)
    # patches: 0 xmin, 1 xmax, 2 ymin, 3 ymax, 4 zmin, 5 zmax
    bcs = [
        make_bc("SUPERSONIC_INLET", p=p, t=T, velocity=(u, 0.0, 0.0)),
        make_bc("SUPERSONIC_OUTLET"),
        make_bc("SLIP_WALL"),
        make_bc("SLIP_WALL"),
        make_bc("SYMMETRY"),
        make_bc("SYMMETRY"),
    ]
    num = make_numerics(False, **{"cfl": 10.0, **num_kw})
    return Case("ramp", mesh, num, bcs, rho, (u, 0.0, 0.0), p, default_linparams(),
                info=dict(mach=mach, T=T, steps=200))


def plate(ni=256, nj=128, nk=64, mach=0.3, re=1.0e5, T=300.0, seed=2024, reorder=1, jitter=0.1, **num_kw):
    a = sound_speed(T)
    u = mach * a
    mu = sutherland(T)
    L = 1.0
    rho = re * mu / (u * L)
    p = rho * R_AIR * T
    mesh = dict(kind=0, ni=ni, nj=nj, nk=nk, x0=-0.25, x1=1.0, y0=0.0, y1=0.25, z0=0.0, z1=0.0625,
                y_ratio=1.08, jitter=jitter, seed=seed, split_ymin=1, split_ymin_x=0.0, reorder=reorder)
    # patches: 0 xmin, 1 xmax, 2 ymin (x > 0: plate), 3 ymax, 4 zmin, 5 zmax, 6 ymin (x < 0)
    bcs = [
        make_bc("FARFIELD", p=p, t=T, velocity=(u, 0.0, 0.0)),
        make_bc("SUBSONIC_OUTLET", p=p),
        make_bc("NO_SLIP_WALL_ADIABATIC"),
        make_bc("FARFIELD", p=p, t=T, velocity=(u, 0.0, 0.0)),
        make_bc("SYMMETRY"),
        make_bc("SYMMETRY"),
        make_bc("SYMMETRY"),
    ]
    num = make_numerics(True, **{"cfl": 10.0, **num_kw})
    return Case("plate", mesh, num, bcs, rho, (u, 0.0, 0.0), p, default_linparams(),
                info=dict(mach=mach, re=re, T=T, mu=mu))


def bump(ni=150, nj=150, nk=150, mach=0.8, re=1.0e6, T=300.0, seed=7, jitter=0.05, **num_kw):
    a = sound_speed(T)
    u = mach * a
    mu = sutherland(T)
    L = 1.0
    rho = re * mu / (u * L)
    p = rho * R_AIR * T
    mesh = dict(kind=1, ni=ni, nj=nj, nk=nk, x0=-1.0, x1=2.0, y0=0.0, y1=1.0, z0=0.0, z1=0.5,
                bump_h=0.1, bump_w=0.25, bump_x=0.5, jitter=jitter, seed=seed, reorder=1)
    bcs = [
        make_bc("FARFIELD", p=p, t=T, velocity=(u, 0.0, 0.0)),
        make_bc("SUBSONIC_OUTLET", p=p),
        make_bc("NO_SLIP_WALL_ADIABATIC"),
        make_bc("FARFIELD", p=p, t=T, velocity=(u, 0.0, 0.0)),
        make_bc("SYMMETRY"),
        make_bc("SYMMETRY"),
    ]
    num = make_numerics(True, **{"cfl": 10.0, **num_kw})
    return Case("bump", mesh, num, bcs, rho, (u, 0.0, 0.0), p, default_linparams(),
                info=dict(mach=mach, re=re, T=T, mu=mu))


def ramp_sa(ni=48, nj=16, nk=4, mach=2.5, re=5.0e6, T=300.0, p=101325.0, **num_kw):
    """The ramp geometry with viscous flow and the Spalart-Allmaras module at
    Re_L = 5e6 (L = 3 m, the ramp length): no-slip ramp and bottom wall,
    supersonic inflow / outflow, slip top wall, symmetry z-faces.  The
    density follows from Re_L; the pressure from the ideal gas at T."""
    a = sound_speed(T)
    u = mach * a
    L = 3.0
    rho = re * sutherland(T) / (u * L)
    p = rho * R_AIR * T
    mesh = dict(kind=0, ni=ni, nj=nj, nk=nk, x0=0.0, x1=3.0, y0=0.0, y1=1.0, z0=0.0, z1=0.25,
                ramp_x=0.5, ramp_deg=15.0)
    bcs = [
        make_bc("SUPERSONIC_INLET", p=p, t=T, velocity=(u, 0.0, 0.0)),
        make_bc("SUPERSONIC_OUTLET"),
        make_bc("NO_SLIP_WALL_ADIABATIC"),
        make_bc("SLIP_WALL"),
        make_bc("SYMMETRY"),
        make_bc("SYMMETRY"),
    ]
    num = make_numerics(True, **{"cfl": 10.0, **num_kw})
    return Case("ramp_sa", mesh, num, bcs, rho, (u, 0.0, 0.0), p, default_linparams(),
                info=dict(mach=mach, re=re, T=T, turbulence="SA"))


def plate_sa(ni=256, nj=128, nk=64, re=5.0e6, **kw):
    """The laminar plate's mesh and conditions at Re_L = 5e6 with the
    Spalart-Allmaras module."""
    c = plate(ni=ni, nj=nj, nk=nk, re=re, **kw)
    c.name = "plate_sa"
    c.info["turbulence"] = "SA"
    return c


def bump_sa(ni=150, nj=150, nk=150, **kw):
    """The transonic tetrahedral bump (Mach 0.8, Re 1e6) with the
    Spalart-Allmaras module."""
    c = bump(ni=ni, nj=nj, nk=nk, **kw)
    c.name = "bump_sa"
    c.info["turbulence"] = "SA"
    return c


def turbulence(case, update=None, nu_inf_ratio=3.0):
    """cfdg.Turbulence block of an SA case (freestream = the case's state)."""
    from . import TURB_POINT_IMPLICIT, Turbulence
    return Turbulence.make(rho_inf=case.rho, p_inf=case.p, update=TURB_POINT_IMPLICIT if update is None else update,
                           nu_inf_ratio=nu_inf_ratio)


CASES = {"ramp": ramp, "plate": plate, "bump": bump, "ramp_sa": ramp_sa, "plate_sa": plate_sa, "bump_sa": bump_sa}
