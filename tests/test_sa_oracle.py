"""Spalart-Allmaras restatement in the oracle, pinned without the reference.

The reference's SA module (spalart_allmaras.hpp) does not compile and is
never dispatched, so no reference output exists for it.  The oracle's SA is
pinned here by what the model's mathematics fixes:

  * the published constants (cw1 = cb1 / kappa^2 + (1 + cb2) / sigma
    = 3.2390678...);
  * closed forms of the closure functions: fv1(chi = cv1) = 1/2,
    fw(r = 1) = 1, fw(r -> 10) -> (1 + cw3^6)^(1/6), Sbar = S and
    P = cb1 S rho nu_tilde far from walls;
  * the eddy viscosity mu_t = rho nu_tilde fv1 = cv1 mu / 2 on a state with
    nu_tilde = cv1 nu everywhere;
  * a manufactured field: nu_tilde linear in space, uniform density,
    temperature and velocity, no walls.  On a uniform hexahedral grid the
    discrete operator is exact for it at cells with no boundary face:
      res = V rho [u . a - (1 + cb2) |a|^2 / sigma - cb1 Sbar_floor nu_tilde]
    (upwind convection of a linear field, the skew-corrected diffusion and the
    cb2 source); a dropped term, a wrong sign (the two sign slips of the
    reference header among them) or a wrong 1/sigma fails it;
  * wall distance: the module's distance is the vertex-stencil relaxation
    that is itself pinned bit-exact to the reference harness.
"""
import numpy as np
import pytest

import cfdg
import cfdg.cases as cases
from oracle_py import Oracle, sa_constants, sa_point


def test_constants_are_the_published_ones():
    k = sa_constants()
    assert k["cb1"] == 0.1355 and k["cb2"] == 0.622 and k["kappa"] == 0.41
    assert k["cw2"] == 0.3 and k["cw3"] == 2.0 and k["cv1"] == 7.1
    assert abs(k["sigma"] - 2.0 / 3.0) < 1e-16
    assert abs(k["cw1"] - 3.2390678) < 1e-7


def test_fv1_half_at_chi_cv1():
    mu, rho = 1.8e-5, 1.2
    o = sa_point(rho, 7.1 * mu / rho, mu, 1.0, 100.0, 0.0)
    assert abs(o["fv1"] - 0.5) < 1e-15


def test_fw_is_one_at_r_one():
    rho, mu, d = 1.2, 1.8e-5, 0.01
    nu = 3.0 * mu / rho  # chi = 3
    kappa = 0.41
    fv1 = 27.0 / (27.0 + 7.1 ** 3)
    fv2 = 1.0 - 3.0 / (1.0 + 3.0 * fv1)
    # r = nu / (Sbar kappa^2 d^2) = 1  <=>  S = nu (1 - fv2) / (kappa^2 d^2)
    S = nu * (1.0 - fv2) / (kappa ** 2 * d ** 2)
    o = sa_point(rho, nu, mu, d, S, 0.0)
    assert abs(o["r"] - 1.0) < 1e-13
    assert abs(o["fw"] - 1.0) < 1e-12
    assert abs(o["D"] - 3.2390678 * rho * nu * nu / d ** 2) < 1e-6 * o["D"]


def test_fw_limit_at_the_r_clip():
    rho, mu = 1.2, 1.8e-5
    o = sa_point(rho, 1e3 * mu / rho, mu, 1e-6, 0.0, 0.0)
    assert o["r"] == 10.0
    assert abs(o["fw"] - 65.0 ** (1.0 / 6.0)) < 1e-9


def test_production_far_from_walls():
    rho, mu, S = 1.1, 1.8e-5, 250.0
    nu = 5.0 * mu / rho
    o = sa_point(rho, nu, mu, 1e8, S, 0.0)
    assert abs(o["sbar"] - S) < 1e-12 * S
    assert abs(o["P"] - 0.1355 * S * rho * nu) < 1e-12 * o["P"]
    assert o["cb2"] == 0.0
    g = sa_point(rho, nu, mu, 1e8, S, 4.0)
    assert abs(g["cb2"] - 0.622 / (2.0 / 3.0) * rho * 4.0) < 1e-12 * g["cb2"]


def _box(ni=10, nj=8, nk=6, u=0.0, bcs_kind="SLIP_WALL"):
    mesh = cfdg.Mesh.generate(kind=0, ni=ni, nj=nj, nk=nk, x0=0.0, x1=1.0, y0=0.0, y1=0.8, z0=0.0, z1=0.6)
    num = cases.make_numerics(True, gradient="LEAST_SQUARES_NODE", limiter="UNLIMITED", cfl=10.0)
    T, p = 300.0, 101325.0
    rho = p / (cases.R_AIR * T)
    if u == 0.0:
        bcs = [cases.make_bc(bcs_kind) for _ in range(6)]
    else:
        bcs = [cases.make_bc("FARFIELD", p=p, t=T, velocity=(u, 0.0, 0.0)), cases.make_bc("SUBSONIC_OUTLET", p=p)] + \
              [cases.make_bc("SYMMETRY") for _ in range(4)]
    return mesh, num, bcs, rho, T, p


def _interior_cells(mesh):
    v = mesh.view
    nf, ni = v.n_faces, v.n_inner_faces
    owner = np.ctypeslib.as_array(v.face_owner, (nf,))
    touched = np.zeros(v.n_cells, bool)
    touched[owner[ni:]] = True
    return np.nonzero(~touched)[0]


@pytest.mark.parametrize("u", [0.0, 40.0])
def test_manufactured_linear_nu_tilde(u):
    mesh, num, bcs, rho, T, p = _box(u=u)
    o = Oracle(mesh, num, bcs)
    o.init_uniform(rho, (u, 0.0, 0.0), p)
    o.set_turbulence(cfdg.Turbulence.make(rho_inf=rho, p_inf=p))
    v = mesh.view
    n = v.n_cells
    cx = np.ctypeslib.as_array(v.cell_cx, (n,))
    cy = np.ctypeslib.as_array(v.cell_cy, (n,))
    cz = np.ctypeslib.as_array(v.cell_cz, (n,))
    mu = cases.sutherland(T)
    a = np.array([3.0, -2.0, 1.5]) * (mu / rho)
    nu = 4.0 * mu / rho + a[0] * cx + a[1] * cy + a[2] * cz
    o.set_turbulence_state(rho * nu)
    o.eval_residual()
    t = o.turbulence()
    inner = _interior_cells(mesh)
    assert len(inner) > 50
    assert np.allclose(t["dist"], 1.0e30)  # no wall patch
    # exact gradient of the linear field
    assert np.allclose(t["grad"][:, inner], a[:, None], rtol=1e-9, atol=1e-12 * np.abs(a).max())
    vol = np.ctypeslib.as_array(v.cell_vol, (n,))[inner]
    expect = vol * rho * (u * a[0] - (1.0 + 0.622) / (2.0 / 3.0) * (a @ a) - 0.1355 * 1e-10 * nu[inner])
    scale = vol * rho * (abs(u * a[0]) + (a @ a))
    err = np.abs(t["res"][inner] - expect) / scale
    assert err.max() < 1e-7, err.max()


def test_eddy_viscosity_at_chi_cv1():
    mesh, num, bcs, rho, T, p = _box()
    o = Oracle(mesh, num, bcs)
    o.init_uniform(rho, (0.0, 0.0, 0.0), p)
    o.set_turbulence(cfdg.Turbulence.make(rho_inf=rho, p_inf=p))
    mu = cases.sutherland(T)
    o.set_turbulence_state(np.full(mesh.n_cells, 7.1 * mu))  # rho nu_tilde = cv1 mu
    o.eval_residual()
    t = o.turbulence()
    assert np.allclose(t["mut"], 0.5 * 7.1 * mu, rtol=1e-13)


def test_freestream_nu_tilde_and_wall_distance():
    case = cases.plate(ni=16, nj=8, nk=4)
    mesh = cfdg.Mesh.generate(**case.mesh)
    o = Oracle(mesh, case.numerics, case.bcs)
    o.init_uniform(case.rho, case.velocity, case.p)
    o.set_turbulence(cfdg.Turbulence.make(rho_inf=case.rho, p_inf=case.p, nu_inf_ratio=3.0))
    o.eval_residual()
    t = o.turbulence()
    T_inf = case.p / (case.rho * cases.R_AIR)
    assert np.allclose(t["rho_nu"], 3.0 * cases.sutherland(T_inf), rtol=1e-13)
    assert np.array_equal(t["dist"], o.wall_distance(stencil=1, max_sweeps=500, rel_tol=1e-8))


def test_point_implicit_update_converges_and_frozen_keeps_state():
    case = cases.plate(ni=16, nj=8, nk=4, re=5.0e6, limiter="UNLIMITED", cfl=50.0)
    mesh = cfdg.Mesh.generate(**case.mesh)
    lp = cases.default_linparams()
    o = Oracle(mesh, case.numerics, case.bcs)
    o.init_uniform(case.rho, case.velocity, case.p)
    o.set_turbulence(cfdg.Turbulence.make(rho_inf=case.rho, p_inf=case.p, update=cfdg.TURB_FROZEN))
    o.implicit_step(lp)
    t = o.turbulence()
    assert np.all(t["rho_nu"] == t["rho_nu"][0])  # the module block is copied unchanged
    o2 = Oracle(mesh, case.numerics, case.bcs)
    o2.init_uniform(case.rho, case.velocity, case.p)
    o2.set_turbulence(cfdg.Turbulence.make(rho_inf=case.rho, p_inf=case.p, update=cfdg.TURB_POINT_IMPLICIT))
    norms = []
    for _ in range(150):
        o2.implicit_step(lp)
        t = o2.turbulence()
        norms.append(np.sqrt(np.mean(t["res"] ** 2)))
    assert np.all(t["rho_nu"] >= 0.0)
    assert norms[-1] < 1e-2 * norms[0], (norms[0], norms[-1])
    assert t["mut"].max() > 0.0


# ---- boundary zones (spalart_allmaras.hpp:50-55, zone_kind :452-468) --------
# An orthogonal box at rest with a uniform nu_tilde = 2.5 nu_inf and one patch
# of each zone: farfield (inflow: ghost nu_inf), subsonic and supersonic outlet
# (outflow: mirror), adiabatic no-slip wall (wall: ghost -nu), slip wall and
# symmetry (symmetry: mirror).  The mass flux is zero on every face, interior
# faces carry no diffusion (uniform field, the skew correction cancels the
# averaged gradient along d || n), so with the Green-Gauss face gradient:
#   V grad_c = sum over c's boundary faces of (nu_g - nu) / 2 A n
#   res_c    = - sum over c's boundary faces of coef (nu_g - nu) / |d| A
#              + V (D - P - cb2 term)(grad_c),     coef = (mu + rho (nu + nu_g) / 2) / sigma,
# |d| = 2 |x_f - x_c|, sources from sa_point (pinned above) at zero strain.

ZONE_BOX_BCS = (("FARFIELD", "inflow"), ("SUBSONIC_OUTLET", "mirror"), ("NO_SLIP_WALL_HEAT_FLUX", "wall"),
                ("SLIP_WALL", "mirror"), ("SYMMETRY", "mirror"), ("SUPERSONIC_OUTLET", "mirror"))


def zone_box():
    T, p = 300.0, 101325.0
    rho = p / (cases.R_AIR * T)
    mesh = cfdg.Mesh.generate(kind=0, ni=6, nj=5, nk=4, x0=0.0, x1=1.2, y0=0.0, y1=0.5, z0=0.0, z1=0.6)
    num = cases.make_numerics(True, gradient="GREEN_GAUSS_FACE", limiter="UNLIMITED", cfl=10.0)
    bcs = []
    for kind, _ in ZONE_BOX_BCS:
        kw = {"p": p, "t": T} if kind in ("FARFIELD", "SUBSONIC_OUTLET") else {}
        bcs.append(cases.make_bc(kind, **kw))
    return mesh, num, bcs, rho, T, p


def zone_box_closed_form(mesh, rho, T, dist, wall_sign=-1.0):
    """(gradient 3 x n, residual n) of the zone box; wall_sign = the wall
    ghost's nu / the cell's (-1 for the model; +1 plants a mirror wall)."""
    A = mesh.arrays()
    nc, nf, ni = mesh.n_cells, mesh.n_faces, mesh.n_inner_faces
    own, pat, area, vol = A["face_owner"], A["face_patch"], A["face_area"], A["cell_vol"]
    n = np.stack([A["face_nx"], A["face_ny"], A["face_nz"]])
    xf = np.stack([A["face_cx"], A["face_cy"], A["face_cz"]])
    xc = np.stack([A["cell_cx"], A["cell_cy"], A["cell_cz"]])
    mu = cases.sutherland(T)
    nu_inf = 3.0 * mu / rho
    nu = 2.5 * nu_inf
    ghost = {"inflow": nu_inf, "mirror": nu, "wall": wall_sign * nu}
    zone = [z for _, z in ZONE_BOX_BCS]
    G = np.zeros((3, nc))
    R = np.zeros(nc)
    for f in range(ni, nf):
        c, nu_g = own[f], ghost[zone[pat[f]]]
        G[:, c] += 0.5 * (nu_g - nu) * area[f] * n[:, f]
        d = 2.0 * np.linalg.norm(xf[:, f] - xc[:, c])
        R[c] -= (mu + rho * 0.5 * (nu + nu_g)) / (2.0 / 3.0) * (nu_g - nu) / d * area[f]
    G /= vol
    for c in range(nc):
        s = sa_point(rho, nu, mu, dist[c], 0.0, G[:, c] @ G[:, c])
        R[c] += vol[c] * (s["D"] - s["P"] - s["cb2"])
    return G, R, nu


def test_boundary_zones_closed_form():
    mesh, num, bcs, rho, T, p = zone_box()
    o = Oracle(mesh, num, bcs)
    o.init_uniform(rho, (0.0, 0.0, 0.0), p)
    o.set_turbulence(cfdg.Turbulence.make(rho_inf=rho, p_inf=p))
    nu = 2.5 * 3.0 * cases.sutherland(T) / rho
    o.set_turbulence_state(np.full(mesh.n_cells, rho * nu))
    o.eval_residual()
    t = o.turbulence()
    assert np.all(t["mdot"] == 0.0)
    G, R, _ = zone_box_closed_form(mesh, rho, T, t["dist"])
    assert np.abs(t["grad"] - G).max() <= 1e-12 * np.abs(G).max()
    assert np.abs(t["res"] - R).max() <= 1e-12 * np.abs(R).max()
    # every zone is touched, and the closed form tells a mirrored wall apart
    own = mesh.array("face_owner")[mesh.n_inner_faces:]
    assert len(set(mesh.array("face_patch")[mesh.n_inner_faces:])) == len(ZONE_BOX_BCS)
    G2, R2, _ = zone_box_closed_form(mesh, rho, T, t["dist"], wall_sign=1.0)
    wall_cells = own[mesh.array("face_patch")[mesh.n_inner_faces:] == 2]
    assert np.abs(R2[wall_cells] - R[wall_cells]).min() > 1e-3 * np.abs(R).max()
