"""Spalart-Allmaras module on the GPU against the CPU oracle.

Cases: flat plate 24x16x8 at Re_L = 5e6 (viscous + SA, MUSCL) with each of
the three gradient methods, the supersonic ramp with no-slip walls at
Re_L = 5e6 (inflow / outflow zones) and the tetrahedral bump; freestream
with 2 % seeded noise in U and 20 % in rho nu_tilde.

Tolerances (FP64 on both sides, same summation order):
  * eddy viscosity, nu_tilde, its gradient, density: 1e-12 relative;
  * face mass flux: 1e-12 of max |mdot|;
  * module residual and point-implicit diagonal: 1e-12 of max |res|, max |diag|;
  * mean-flow residual with mu_t: 1e-12 of the per-cell flux scale; spectral
    radius with the module's contribution, dt and alpha: 1e-12 relative;
  * Jacobian blocks with mu_t: 1e-12 of max |B|;
  * wall distance: bit-identical (same relaxation, same order);
  * five implicit steps (frozen and point-implicit module): state and rho nu_tilde
    within 1e-9 of each variable's maximum (the BiCGSTAB dot products are tree
    reductions, 1e-9 per solve as in the mean-flow step test).
"""
import numpy as np
import pytest

import cfdg
import cfdg.cases as cases
from helpers import assert_close_scaled, perturbed_state, rel_err
from oracle_py import Oracle

pytestmark = pytest.mark.gpu


def _case(gradient="LEAST_SQUARES_NODE"):
    return cases.plate(ni=24, nj=16, nk=8, re=5.0e6, gradient=gradient)


def _pair(case, update=cfdg.TURB_POINT_IMPLICIT, seed=3):
    mesh = cfdg.Mesh.generate(**case.mesh)
    s = cfdg.Solver(mesh, case.numerics, case.bcs)
    o = Oracle(mesh, case.numerics, case.bcs)
    U = perturbed_state(s, case, seed=seed)
    turb = cfdg.Turbulence.make(rho_inf=case.rho, p_inf=case.p, update=update)
    s.set_turbulence(turb)
    o.set_turbulence(turb)
    rnu0 = s.turbulence()["rho_nu"]
    rng = np.random.default_rng(seed + 100)
    rnu = rnu0 * (1.0 + 0.2 * rng.standard_normal(rnu0.shape))
    for x in (s, o):
        x.set_state(U)
        x.set_turbulence_state(rnu)
    return mesh, s, o


SA_GEOMETRIES = {
    # supersonic ramp: inflow and outflow zones, the freestream override of inflowing faces
    "ramp_sa": lambda: cases.ramp_sa(),
    # Kuhn tetrahedra (gather gradient kernel, 70-neighbour stencils)
    "bump_sa": lambda: cases.bump_sa(ni=8, nj=8, nk=8),
}


@pytest.mark.parametrize("gradient", ["LEAST_SQUARES_NODE", "GREEN_GAUSS_FACE", "GREEN_GAUSS_CELL"])
def test_sa_residual_matches_oracle(gradient):
    _check_residual(_case(gradient))


@pytest.mark.parametrize("name", sorted(SA_GEOMETRIES))
def test_sa_residual_matches_oracle_other_geometries(name):
    _check_residual(SA_GEOMETRIES[name]())


def _check_residual(case):
    mesh, s, o = _pair(case)
    s.eval_residual()
    o.eval_residual()
    tg, to = s.turbulence(), o.turbulence()
    assert np.array_equal(tg["dist"], to["dist"])
    for k in ("nu", "mut"):
        assert rel_err(tg[k], to[k]).max() <= 1e-12, k
    assert_close_scaled(tg["grad"], to["grad"], np.abs(to["grad"]).max(), 1e-12, "nu_tilde gradient")
    assert_close_scaled(tg["mdot"], to["mdot"], np.abs(to["mdot"]).max(), 1e-12, "mass flux")
    assert_close_scaled(tg["res"], to["res"], np.abs(to["res"]).max(), 1e-12, "module residual")
    assert_close_scaled(tg["diag"], to["diag"], np.abs(to["diag"]).max(), 1e-12, "module diagonal")
    Rg, lg = s.residual()
    Ro, lo = o.residual()
    assert_close_scaled(Rg, Ro, o.flux_scale(), 1e-12, "mean residual with mu_t")
    assert rel_err(lg, lo).max() <= 1e-12


def test_sa_jacobian_and_dt_match_oracle():
    mesh, s, o = _pair(_case())
    for x in (s, o):
        x.eval_residual()
        x.compute_dt()
        x.assemble_jacobian()
    (dtg, ag), (dto, ao) = s.dt(), o.dt()
    assert rel_err(dtg, dto).max() <= 1e-12 and rel_err(ag, ao).max() <= 1e-12
    Bg, Bo = s.bsr_values(), o.bsr_values()
    assert_close_scaled(Bg, Bo, np.abs(Bo).max(), 1e-12, "Jacobian blocks with mu_t")


@pytest.mark.parametrize("update", [cfdg.TURB_POINT_IMPLICIT, cfdg.TURB_FROZEN])
def test_sa_implicit_steps_match_oracle(update):
    mesh, s, o = _pair(_case(), update=update)
    lp = cases.default_linparams()
    for _ in range(5):
        s.implicit_step(lp)
        o.implicit_step(lp)
    Ug, Uo = s.get_state(), o.get_state()
    assert_close_scaled(Ug, Uo, np.abs(Uo).max(axis=1, keepdims=True), 1e-9, "state after 5 steps")
    rg, ro = s.turbulence()["rho_nu"], o.turbulence()["rho_nu"]
    assert np.abs(rg - ro).max() <= 1e-9 * np.abs(ro).max()
    assert np.all(rg >= 0.0)


def _sa_plate_case():
    return cases.plate(ni=16, nj=8, nk=4, re=5.0e6, limiter="UNLIMITED", cfl=1000.0)


@pytest.fixture(scope="module")
def sa_oracle_converged():
    case = _sa_plate_case()
    mesh = cfdg.Mesh.generate(**case.mesh)
    o = Oracle(mesh, case.numerics, case.bcs)
    U0 = perturbed_state(o, case)
    turb = cfdg.Turbulence.make(rho_inf=case.rho, p_inf=case.p)
    o.set_turbulence(turb)
    o.set_state(U0)
    lp = cases.default_linparams()
    l0 = None
    for _ in range(8000):
        l2 = np.array(o.implicit_step(lp).l2)
        l0 = l2 if l0 is None else l0
        if np.max(l2 / l0) <= 1e-11:
            break
    else:
        pytest.fail("oracle did not converge")
    o.eval_residual()
    return mesh, U0, turb, o


@pytest.mark.parametrize("form", ["fp64", "fp32", "bf16"])
def test_sa_steady_state_matches_oracle(sa_oracle_converged, form):
    """Plate 16x8x4, Re_L 5e6, unlimited MUSCL, CFL 1000, point-implicit SA,
    from 2 % noise: the FP64 oracle and the device in each precision form
    (FP64; FP32 blocks + FP32 Krylov; BF16 blocks + FP32 Krylov, both inside
    the FP64 defect correction) run to a relative residual of 1e-11; per
    field, rho, u, v, w, p within 1e-8 of their freestream scales, rho
    nu_tilde within 1e-8 of its freestream value, wall force within 1e-8
    relative."""
    case = _sa_plate_case()
    mesh, U0, turb, o = sa_oracle_converged
    lp = cases.default_linparams()
    s = cfdg.Solver(mesh, case.numerics, case.bcs)
    s.init_uniform(case.rho, case.velocity, case.p)
    s.set_turbulence(turb)
    s.set_state(U0)
    if form == "fp32":
        s.set_precision(cfdg.Precision.make(cfdg.PREC_FP32, cfdg.PREC_FP32, max_outer=8, inner_rel_tol=0.1))
    elif form == "bf16":
        s.set_precision(cfdg.Precision.make(cfdg.PREC_BF16, cfdg.PREC_FP32, max_outer=16, inner_rel_tol=0.1))
    r = s.run_steady(lp, 8000, rel_tol=1e-11)
    assert r.status == cfdg.CONVERGED, (r.status, r.steps, r.rel)
    s.eval_residual()

    def prim(U):
        rho = U[0]
        u, v, w = U[1] / rho, U[2] / rho, U[3] / rho
        return np.stack([rho, u, v, w, (cases.GAMMA - 1.0) * (U[4] - 0.5 * rho * (u * u + v * v + w * w))])

    speed = float(np.linalg.norm(case.velocity))
    scale = np.array([case.rho, speed, speed, speed, case.p])[:, None]
    err = np.max(np.abs(prim(s.get_state()) - prim(o.get_state())) / scale, axis=1)
    assert np.all(err <= 1e-8), dict(zip(("rho", "u", "v", "w", "p"), err))
    tg, to = s.turbulence(), o.turbulence()
    rnu_inf = to["rho_nu"].max()
    assert np.abs(tg["rho_nu"] - to["rho_nu"]).max() <= 1e-8 * rnu_inf
    walls = [p for p, b in enumerate(case.bcs) if b.type in (cfdg.BC["NO_SLIP_WALL"], cfdg.BC["NO_SLIP_WALL_HEAT_FLUX"])]
    fg = s.boundary_integrals()[2][walls].sum(axis=0)
    fo = o.boundary_integrals()[2][walls].sum(axis=0)
    assert np.linalg.norm(fg - fo) <= 1e-8 * np.linalg.norm(fo), (fg, fo)


def test_boundary_zones_closed_form_on_device():
    """The device's zone ghosts (wall -nu, inflow nu_inf, mirror elsewhere),
    Green-Gauss nu_tilde gradient and boundary-face diffusion on the zone box
    of tests/test_sa_oracle.py, against its closed form: 1e-12 of the maxima."""
    from test_sa_oracle import zone_box, zone_box_closed_form
    mesh, num, bcs, rho, T, p = zone_box()
    s = cfdg.Solver(mesh, num, bcs)
    s.init_uniform(rho, (0.0, 0.0, 0.0), p)
    s.set_turbulence(cfdg.Turbulence.make(rho_inf=rho, p_inf=p))
    nu = 2.5 * 3.0 * cases.sutherland(T) / rho
    s.set_turbulence_state(np.full(mesh.n_cells, rho * nu))
    s.eval_residual()
    t = s.turbulence()
    G, R, _ = zone_box_closed_form(mesh, rho, T, t["dist"])
    assert np.abs(t["grad"] - G).max() <= 1e-12 * np.abs(G).max()
    assert np.abs(t["res"] - R).max() <= 1e-12 * np.abs(R).max()
