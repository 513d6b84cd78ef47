"""Forward Euler and dual time stepping (BDF1, BDF2) on the GPU against the
reference's own Solver runs (tests/golden/time_*.npz) and the CPU oracle.

Tolerances:
  * forward Euler: the residual is evaluated with the reference's operation
    order (1e-12 per evaluation, test_gpu_parity); after 25 steps the state
    is within 1e-11 of each variable's scale (density, the largest momentum
    component for all three, energy); the residual norms (device
    tree reductions of squares, near zero close to the uniform start) within
    1e-9 relative or 1e-12 of the largest norm;
  * dual time: every sub-iteration is an implicit step whose BiCGSTAB dot
    products are tree reductions on the device (1e-9 per solve, as in the
    implicit step tests); the sub-iteration counts equal the oracle's and the
    state is within 1e-9 of the same scales.
"""
import numpy as np
import pytest

import cfdg
from golden_io import bcs_of, linparams_of, load, mesh_of, numerics_of
from helpers import assert_close_scaled
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

NAMES = ["time_ramp", "time_plate"]


def _scale(ref):
    """Per-variable scale of a conservative state: rho, one momentum scale for
    the three components (the freestream has v = w = 0), rho E."""
    m = np.abs(ref).max(axis=1)
    return np.array([m[0], m[1:4].max(), m[1:4].max(), m[1:4].max(), m[4]])[:, None]


def _pair(fx, cfl=None):
    num = numerics_of(fx)
    if cfl is not None:
        num.cfl = cfl
    mesh = mesh_of(fx)
    s = cfdg.Solver(mesh, num, bcs_of(fx))
    o = Oracle(mesh, num, bcs_of(fx))
    rho, u, v, w, p = fx["init"]
    for x in (s, o):
        x.init_uniform(rho, (u, v, w), p)
    return s, o


@pytest.mark.parametrize("name", NAMES)
def test_forward_euler_matches_reference(name):
    fx = load(name)
    cfl, n = fx["explicit"]
    s, o = _pair(fx, cfl=float(cfl))
    for _ in range(int(n)):
        st = s.explicit_step()
        l2o = o.explicit_step()
        assert st.linear.iterations == 0
        np.testing.assert_allclose(st.l2, l2o, rtol=1e-9, atol=1e-12 * float(np.max(l2o)))
    ref = fx["ref_explicit_U"].reshape(5, -1)
    U = s.get_state()
    assert_close_scaled(U, ref, _scale(ref), 1e-11, "forward Euler state")
    assert_close_scaled(U, o.get_state(), _scale(ref), 1e-11, "forward Euler vs oracle")


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("order", [1, 2])
def test_dual_time_matches_reference(name, order):
    fx = load(name)
    dt, n_phys, n_sub, tol = fx["unsteady"]
    lp = linparams_of(fx)
    s, o = _pair(fx)
    s.set_unsteady(dt, order)
    o.set_unsteady(dt, order)
    sub_g = s.run_unsteady(lp, int(n_phys), int(n_sub), tol)
    sub_o = o.run_unsteady(lp, int(n_phys), int(n_sub), tol)
    assert np.array_equal(sub_g, sub_o)
    ref = fx[f"ref_bdf{order}_U"].reshape(5, -1)
    scale = _scale(ref)
    assert_close_scaled(s.get_state(), ref, scale, 1e-9, f"BDF{order} state vs reference")
    assert_close_scaled(s.get_state(), o.get_state(), scale, 1e-9, f"BDF{order} state vs oracle")


def test_dual_time_residual_and_shift_match_oracle():
    """One BDF2 physical step on the viscous plate from a noisy state: the
    residual with the pseudo-source and the shifted diagonal within 1e-12."""
    fx = load("time_plate")
    dt = float(fx["unsteady"][0])
    s, o = _pair(fx)
    rng = np.random.default_rng(5)
    U = s.get_state() * (1.0 + 0.01 * rng.standard_normal((5, s.n_cells)))
    for x in (s, o):
        x.set_state(U)
        x.set_unsteady(dt, 2)
        x.set_state(U * (1.0 + 1e-3))  # U differs from U^n = U^{n-1}
        x.eval_residual()
        x.compute_dt()
    Rg, _ = s.residual()
    Ro, _ = o.residual()
    assert_close_scaled(Rg, Ro, o.flux_scale() + np.abs(Ro), 1e-12, "residual with the BDF pseudo-source")
    (dtg, ag), (dto, ao) = s.dt(), o.dt()
    np.testing.assert_allclose(ag, ao, rtol=1e-12, atol=0.0)


def test_bdf2_differs_from_bdf1_after_first_step():
    fx = load("time_ramp")
    dt = float(fx["unsteady"][0])
    lp = linparams_of(fx)
    a, _ = _pair(fx)
    b, _ = _pair(fx)
    a.set_unsteady(dt, 1)
    b.set_unsteady(dt, 2)
    a.run_unsteady(lp, 1, 6, 1e-3)
    b.run_unsteady(lp, 1, 6, 1e-3)
    assert np.array_equal(a.get_state(), b.get_state())
    a.run_unsteady(lp, 1, 6, 1e-3)
    b.run_unsteady(lp, 1, 6, 1e-3)
    assert not np.array_equal(a.get_state(), b.get_state())


def test_time_stepping_errors():
    fx = load("time_plate")
    s, _ = _pair(fx)
    rho, p = float(fx["init"][0]), float(fx["init"][4])
    with pytest.raises(RuntimeError):
        s.set_unsteady(0.0, 2)
    with pytest.raises(RuntimeError):
        s.set_unsteady(1e-4, 3)
    with pytest.raises(RuntimeError):
        s.run_unsteady(linparams_of(fx), 1, 5, 1e-3)  # not enabled
    s.set_unsteady(1e-4, 2)
    with pytest.raises(RuntimeError):
        s.explicit_step()
    with pytest.raises(RuntimeError):
        s.set_turbulence(cfdg.Turbulence.make(rho_inf=rho, p_inf=p))
    s.set_unsteady(0.0, 0)
    s.set_turbulence(cfdg.Turbulence.make(rho_inf=rho, p_inf=p))
    with pytest.raises(RuntimeError):
        s.explicit_step()
    with pytest.raises(RuntimeError):
        s.set_unsteady(1e-4, 1)


@pytest.mark.parametrize("geometry", ["plate_sa", "bump_sa"])
def test_dual_time_with_frozen_turbulence_matches_oracle(geometry):
    """BDF2 dual time stepping with the Spalart-Allmaras module frozen (the
    module block copied unchanged by the implicit step, implicit_euler.hpp:
    64-72): eddy viscosity in the mean-flow residual and Jacobian, the
    module's spectral radius in the local step, the module state unchanged.
    Sub-iteration counts equal the oracle's; state within 1e-9 of the scales;
    rho nu_tilde unchanged on both sides.  The point-implicit update has no
    pseudo-time source and stays refused."""
    import cfdg.cases as cases
    from helpers import perturbed_state

    case = cases.plate_sa(ni=12, nj=8, nk=4) if geometry == "plate_sa" else cases.bump_sa(ni=5, nj=5, nk=5)
    mesh = cfdg.Mesh.generate(**case.mesh)
    s = cfdg.Solver(mesh, case.numerics, case.bcs)
    o = Oracle(mesh, case.numerics, case.bcs)
    U = perturbed_state(s, case, seed=9)
    frozen = cases.turbulence(case, update=cfdg.TURB_FROZEN)
    for x in (s, o):
        x.init_uniform(case.rho, case.velocity, case.p)
        x.set_state(U)
        x.set_turbulence(frozen)
    rnu0 = s.turbulence()["rho_nu"]
    speed = float(np.linalg.norm(case.velocity))
    dt = 0.5 * float(np.cbrt(mesh.array("cell_vol").min())) / speed
    for x in (s, o):
        x.set_unsteady(dt, 2)
    lp = case.linparams
    sub_g = s.run_unsteady(lp, 3, 20, 1e-6)
    sub_o = o.run_unsteady(lp, 3, 20, 1e-6)
    assert np.array_equal(sub_g, sub_o), (sub_g, sub_o)
    Uo = o.get_state()
    assert_close_scaled(s.get_state(), Uo, _scale(Uo), 1e-9, f"{geometry} BDF2 with frozen SA vs oracle")
    assert np.array_equal(s.turbulence()["rho_nu"], rnu0)
    assert np.array_equal(o.turbulence()["rho_nu"], rnu0)
    t = cfdg.Solver(mesh, case.numerics, case.bcs)
    t.init_uniform(case.rho, case.velocity, case.p)
    t.set_turbulence(cases.turbulence(case, update=cfdg.TURB_POINT_IMPLICIT))
    with pytest.raises(RuntimeError):
        t.set_unsteady(dt, 2)
