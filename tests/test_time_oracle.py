"""Forward Euler and dual time stepping (BDF1, BDF2) of the CPU oracle against
the reference's own Solver runs (tests/golden/time_*.npz, written by
tests/golden/make_golden.py from the compiled reference: ForwardEuler in
steady mode, BackwardEuler in UnsteadyMode<1> / UnsteadyMode<2>).

Bit-exact equality is required, as for the implicit steps."""
import numpy as np
import pytest

from golden_io import bcs_of, linparams_of, load, mesh_of, numerics_of
from oracle_py import Oracle

NAMES = ["time_ramp", "time_plate"]


def _oracle(fx, cfl=None):
    num = numerics_of(fx)
    if cfl is not None:
        num.cfl = cfl
    o = Oracle(mesh_of(fx), num, bcs_of(fx))
    rho, u, v, w, p = fx["init"]
    o.init_uniform(rho, (u, v, w), p)
    return o


@pytest.mark.parametrize("name", NAMES)
def test_forward_euler_bit_exact(name):
    fx = load(name)
    cfl, n = fx["explicit"]
    o = _oracle(fx, cfl=float(cfl))
    U0 = o.get_state().copy()
    for _ in range(int(n)):
        o.explicit_step()
    U = o.get_state()
    assert not np.array_equal(U, U0)
    assert np.array_equal(U.ravel(), fx["ref_explicit_U"])


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("order", [1, 2])
def test_dual_time_bit_exact(name, order):
    fx = load(name)
    dt, n_phys, n_sub, tol = fx["unsteady"]
    o = _oracle(fx)
    o.set_unsteady(dt, order)
    sub = o.run_unsteady(linparams_of(fx), int(n_phys), int(n_sub), tol)
    assert np.all(sub >= 1)
    assert np.array_equal(o.get_state().ravel(), fx[f"ref_bdf{order}_U"])


def test_bdf_orders_differ_and_unsteady_source_vanishes_at_rest():
    """BDF2 differs from BDF1 after the first physical step (the first step is
    BDF1 in both); at U = U^n = U^{n-1} the pseudo-source is exactly zero, so
    the residual equals the steady one."""
    fx = load("time_ramp")
    dt = float(fx["unsteady"][0])
    lp = linparams_of(fx)
    a, b = _oracle(fx), _oracle(fx)
    a.set_unsteady(dt, 1)
    b.set_unsteady(dt, 2)
    a.run_unsteady(lp, 1, 6, 1e-3)
    b.run_unsteady(lp, 1, 6, 1e-3)
    assert np.array_equal(a.get_state(), b.get_state())
    a.run_unsteady(lp, 1, 6, 1e-3)
    b.run_unsteady(lp, 1, 6, 1e-3)
    assert not np.array_equal(a.get_state(), b.get_state())
    s, u = _oracle(fx), _oracle(fx)
    u.set_unsteady(dt, 2)
    s.eval_residual()
    u.eval_residual()
    assert np.array_equal(s.residual()[0], u.residual()[0])


def test_invalid_time_settings_rejected():
    fx = load("time_ramp")
    o = _oracle(fx)
    with pytest.raises(RuntimeError):
        o.set_unsteady(0.0, 2)
    with pytest.raises(RuntimeError):
        o.set_unsteady(1e-4, 3)
    o.set_unsteady(1e-4, 2)
    with pytest.raises(RuntimeError):
        o.explicit_step()
