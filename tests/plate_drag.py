"""Skin-friction drag of the laminar flat plate of the baseline (Mach 0.3,
Re_L = 1e5, adiabatic no-slip plate from x = 0 to 1) against the Blasius
value C_D = 1.328 / sqrt(Re_L) for one side of the plate: a sanity check of
the viscous flux, the wall boundary condition and the boundary integrals
(residual_kernel.hpp:96-169) at a converged state.  A 64x32x2 plate with the
wall-normal stretching of the baseline, unlimited MUSCL (the limited scheme
stalls on this plate), CFL 1000; the drag is the x-force on the plate patch
over (rho u^2 / 2) L W.  Shared by the oracle and the device test."""
import math

import numpy as np

import cfdg
import cfdg.cases as cases

RE_L = 1.0e5
STEPS = 2250


def case():
    return cases.plate(ni=64, nj=32, nk=2, limiter="UNLIMITED", cfl=1000.0, jitter=0.0, re=RE_L)


def drag_history(s, c):
    """The drag coefficient after steps 2000 and 2250 (steady to 1e-3 by step
    2000 on this mesh)."""
    s.init_uniform(c.rho, c.velocity, c.p)
    lp = cases.default_linparams(order=cfdg.SGS_SEQUENTIAL)
    q = 0.5 * c.rho * c.velocity[0] ** 2
    area = 1.0 * (c.mesh["z1"] - c.mesh["z0"])
    out = []
    for k in range(1, STEPS + 1):
        s.implicit_step(lp)
        if k in (2000, STEPS):
            s.eval_residual()  # the integrals of the current state
            force = s.boundary_integrals()[2]
            out.append(force[2][0] / (q * area))  # patch 2: the plate (y = 0, x > 0)
    return out


def check(cd):
    blasius = 1.328 / math.sqrt(RE_L)
    assert abs(cd[-1] - cd[-2]) < 1e-3 * cd[-1]  # steady
    assert abs(cd[-1] / blasius - 1.0) < 0.05, (cd[-1], blasius)
