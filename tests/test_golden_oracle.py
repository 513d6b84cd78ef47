"""The CPU oracle against the reference's own outputs (tests/golden/*.npz,
written by tests/golden/make_golden.py from the compiled reference).

Bit-exact equality is required: the oracle restates every reference routine
with the same operation order and FP contraction disabled."""
import numpy as np
import pytest

from golden_io import bcs_of, fixture_names, linparams_of, load, mesh_of, numerics_of
from oracle_py import Oracle

NAMES = fixture_names()


def _run_oracle(fx):
    mesh = mesh_of(fx)
    o = Oracle(mesh, numerics_of(fx), bcs_of(fx))
    o.set_state(fx["U"])
    o.eval_residual()
    return mesh, o


@pytest.mark.parametrize("name", NAMES)
def test_residual_pipeline_bit_exact(name):
    fx = load(name)
    mesh, o = _run_oracle(fx)
    q = o.primitives()
    g, phi = o.gradients()
    R, lam = o.residual()
    assert np.array_equal(q.ravel(), fx["ref_q"])
    assert np.array_equal(g.ravel(), fx["ref_grad"])
    assert np.array_equal(phi.ravel(), fx["ref_phi"])
    assert np.array_equal(R.ravel(), fx["ref_res"])
    assert np.array_equal(lam, fx["ref_lam"])


@pytest.mark.parametrize("name", NAMES)
def test_jacobian_and_linear_solve_bit_exact(name):
    fx = load(name)
    mesh, o = _run_oracle(fx)
    o.compute_dt()
    dt, alpha = o.dt()
    assert np.array_equal(dt, fx["ref_dt"])
    assert np.array_equal(alpha, fx["ref_alpha"])
    o.assemble_jacobian()
    rp, cols, dg = o.bsr_pattern()
    assert np.array_equal(rp, fx["ref_row_ptr"])
    assert np.array_equal(cols, fx["ref_cols"])
    assert np.array_equal(dg, fx["ref_diag"])
    assert np.array_equal(o.bsr_values().ravel(), fx["ref_values"])
    du, res = o.linear_solve(linparams_of(fx))
    assert [res.status, res.iterations] == list(fx["ref_lin_status"])
    assert np.array_equal(du, fx["ref_du"])
    assert [res.initial_residual, res.final_residual] == list(fx["ref_lin_norms"])


@pytest.mark.parametrize("name", NAMES)
def test_implicit_steps_bit_exact(name):
    fx = load(name)
    mesh = mesh_of(fx)
    o = Oracle(mesh, numerics_of(fx), bcs_of(fx))
    rho, u, v, w, p = fx["init"]
    o.init_uniform(rho, (u, v, w), p)
    lp = linparams_of(fx)
    for _ in range(int(fx["steps"][0])):
        o.implicit_step(lp)
    assert np.array_equal(o.get_state().ravel(), fx["ref_steps_U"])


@pytest.mark.parametrize("name", NAMES)
def test_boundary_integrals(name):
    """Per-patch mass and energy fluxes equal the reference's
    ResidualKernel::boundary_integrals bit for bit.  The force (momentum flux
    per patch) has no reference counterpart; it is pinned by telescoping: the
    interior face fluxes cancel in sum_c R_c, so sum over patches of each
    boundary integral equals the sum of the cell residuals of that variable."""
    fx = load(name)
    mesh, o = _run_oracle(fx)
    mass, energy, force = o.boundary_integrals()
    assert np.array_equal(mass, fx["ref_bnd_mass"])
    assert np.array_equal(energy, fx["ref_bnd_energy"])
    R, _ = o.residual()
    S = o.flux_scale()
    for k, total in ((0, mass.sum()), (1, force[:, 0].sum()), (2, force[:, 1].sum()), (3, force[:, 2].sum()),
                     (4, energy.sum())):
        assert abs(total - R[k].sum()) <= 1e-12 * S[k].sum(), (k, total, R[k].sum())
