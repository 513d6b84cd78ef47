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


@pytest.mark.parametrize("name", NAMES)
def test_wall_distance_bit_exact(name):
    """Wall distance from the no-slip patches equals the reference's
    MeshAuxGeometry::add_wall_distance bit for bit, with the face-neighbour
    stencil (100 sweeps, 1e-4) and the vertex-neighbour one (500, 1e-8)."""
    fx = load(name)
    mesh, o = _run_oracle(fx)
    assert np.array_equal(o.wall_distance(0, 100, 1e-4), fx["ref_wall_dist_face"])
    assert np.array_equal(o.wall_distance(1, 500, 1e-8), fx["ref_wall_dist_node"])


def test_wall_distance_closed_form():
    """On an undistorted hex box with a no-slip wall at y = 0 the face-stencil
    graph distance of a cell is exactly its centroid height (straight column
    of cells), a closed form independent of the reference."""
    import cfdg
    import cfdg.cases as cases
    case = cases.plate(ni=6, nj=5, nk=2, jitter=0.0, reorder=0)
    mesh = cfdg.Mesh.generate(**case.mesh)
    bcs = list(case.bcs)
    for b in bcs:
        b.type = 7  # symmetry everywhere ...
    bcs[2].type = 5  # ... but a no-slip wall on ymin (x >= 0)
    bcs[6].type = 5  # and on ymin (x < 0)
    o = Oracle(mesh, case.numerics, bcs)
    d = o.wall_distance(0, 100, 1e-12)
    cy = mesh.array("cell_cy")
    assert np.allclose(d, cy, rtol=0, atol=1e-14)
