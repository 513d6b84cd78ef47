"""Supersonic flow over a 15 degree wedge at Mach 2.5: the attached oblique
shock's downstream state from the theta-beta-Mach relation and the
Rankine-Hugoniot jump (gamma = 1.4): beta = 36.945 deg, p2/p1 = 2.4675,
rho2/rho1 = 1.8665, M2 = 1.8735, flow turned by 15 deg.  The baseline's
inviscid ramp (48x16 cells on [0,3]x[0,1]) is the same wedge in a channel
whose ramp narrows it to a third, which chokes the Mach 2.5 inflow; this case
keeps the ramp geometry and numerics (HLLC, MUSCL, node LSQ, Venkatakrishnan,
implicit with BiCGSTAB + SGS) on [0,2]x[0,2], where the shock leaves through
the outflow before it reaches the top wall.  Shared by the oracle test and the
device test."""
import math

import numpy as np

import cfdg
import cfdg.cases as cases

GAMMA = 1.4
THETA = 15.0
MACH = 2.5


def theory():
    """(beta, p2/p1, rho2/rho1, M2) from the theta-beta-Mach relation
    (weak branch, bisection) and the normal-shock relations."""
    th = math.radians(THETA)

    def f(b):
        return math.tan(th) - 2.0 / math.tan(b) * (MACH ** 2 * math.sin(b) ** 2 - 1.0) / (
            MACH ** 2 * (GAMMA + math.cos(2.0 * b)) + 2.0)
    lo, hi = math.asin(1.0 / MACH) + 1e-9, math.radians(60.0)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if (f(lo) > 0) == (f(mid) > 0):
            lo = mid
        else:
            hi = mid
    b = 0.5 * (lo + hi)
    mn = MACH * math.sin(b)
    p = 1.0 + 2.0 * GAMMA / (GAMMA + 1.0) * (mn ** 2 - 1.0)
    r = (GAMMA + 1.0) * mn ** 2 / ((GAMMA - 1.0) * mn ** 2 + 2.0)
    mn2 = math.sqrt((1.0 + 0.5 * (GAMMA - 1.0) * mn ** 2) / (GAMMA * mn ** 2 - 0.5 * (GAMMA - 1.0)))
    return math.degrees(b), p, r, mn2 / math.sin(b - th)


def case():
    c = cases.ramp(ni=48, nj=48, nk=1, mach=MACH, cfl=20.0)
    c.mesh.update(x1=2.0, y1=2.0, z1=0.05)
    return c


STEPS = 300


def run(s, c):
    s.init_uniform(c.rho, c.velocity, c.p)
    lp = cases.default_linparams(order=cfdg.SGS_SEQUENTIAL)
    for _ in range(STEPS):
        s.implicit_step(lp)
    return s.get_state()


def check(mesh, c, U):
    beta, p21, r21, m2 = theory()
    assert abs(beta - 36.945) < 1e-3 and abs(p21 - 2.4675) < 1e-4 and abs(r21 - 1.8665) < 1e-4 and abs(m2 - 1.8735) < 1e-4
    rho = U[0]
    u, v, w = U[1] / rho, U[2] / rho, U[3] / rho
    p = (GAMMA - 1.0) * (U[4] - 0.5 * rho * (u * u + v * v + w * w))
    mach = np.sqrt(u * u + v * v + w * w) / np.sqrt(GAMMA * p / rho)
    cx, cy = mesh.array("cell_cx"), mesh.array("cell_cy")
    wall = (cx - 0.5) * math.tan(math.radians(THETA))
    shock = (cx - 0.5) * math.tan(math.radians(beta))
    # between the wedge and the shock, two cells clear of the shock's smearing
    zone = (cx > 1.0) & (cx < 1.8) & (cy > wall + 0.05) & (cy < shock - 0.2)
    assert zone.sum() > 50
    assert abs(np.mean(p[zone]) / c.p / p21 - 1.0) < 2e-3
    assert abs(np.mean(rho[zone]) / c.rho / r21 - 1.0) < 2e-3
    assert abs(np.mean(mach[zone]) / m2 - 1.0) < 3e-3
    assert abs(np.mean(np.degrees(np.arctan2(v[zone], u[zone]))) - THETA) < 0.02
    assert np.max(np.abs(p[zone] / c.p / p21 - 1.0)) < 5e-3
    # the inflow ahead of the wedge is untouched
    up = cx < 0.4
    assert np.max(np.abs(p[up] / c.p - 1.0)) < 1e-6
