"""HIP path with a prescribed eddy-viscosity field (cfdg_set_eddy_viscosity)
against the reference's own outputs (tests/golden/plate_mut.npz: the
reference's ResidualKernel::apply / JacobianKernel::apply with that mut) and
against the oracle on a mesh of several tiles, so that faces on tile
boundaries take mu_t of the halo side from device memory.

Tolerances: residual |dR| <= 1e-12 * sum_f |F| per cell and variable;
spectral radius 1e-12 relative; Jacobian blocks max|dB| <= 1e-12 max|B| per
block; boundary mass and energy 1e-12 of the summed |F|."""
import numpy as np
import pytest

import cfdg
from golden_io import bcs_of, load, mesh_of, numerics_of
from helpers import assert_close_scaled, perturbed_state, small_cases
from oracle_py import Oracle

pytestmark = pytest.mark.gpu


def _check_against(gpu, ora, what):
    Rg, lg = gpu.residual()
    Ro, lo = ora.residual()
    S = ora.flux_scale()
    assert_close_scaled(Rg, Ro, S, 1e-12, f"{what}: residual")
    assert_close_scaled(lg, lo, np.abs(lo), 1e-12, f"{what}: spectral radius")
    for s in (gpu, ora):
        s.compute_dt()
        s.assemble_jacobian()
    Bg, Bo = gpu.bsr_values(), ora.bsr_values()
    assert_close_scaled(Bg, Bo, np.abs(Bo).max(axis=(1, 2), keepdims=True), 1e-12, f"{what}: Jacobian blocks")
    return Rg, lg, Bg


def test_mut_matches_reference_golden():
    fx = load("plate_mut")
    mesh = mesh_of(fx)
    gpu = cfdg.Solver(mesh, numerics_of(fx), bcs_of(fx))
    gpu.set_state(fx["U"])
    gpu.set_eddy_viscosity(fx["mut"])
    gpu.eval_residual()
    ora = Oracle(mesh, numerics_of(fx), bcs_of(fx))  # flux scale of the tolerance only
    ora.set_state(fx["U"])
    ora.set_eddy_viscosity(fx["mut"])
    ora.eval_residual()
    S = ora.flux_scale()
    Rg, lg = gpu.residual()
    assert_close_scaled(Rg.ravel(), fx["ref_res"], S.ravel(), 1e-12, "residual vs reference")
    assert_close_scaled(lg, fx["ref_lam"], np.abs(fx["ref_lam"]), 1e-12, "spectral radius vs reference")
    mass, energy, _ = gpu.boundary_integrals()
    assert_close_scaled(mass, fx["ref_bnd_mass"], S[0].sum(), 1e-12, "boundary mass")
    assert_close_scaled(energy, fx["ref_bnd_energy"], S[4].sum(), 1e-12, "boundary energy")
    gpu.compute_dt()
    _, alpha = gpu.dt()
    assert_close_scaled(alpha, fx["jac_alpha"], np.abs(fx["jac_alpha"]), 1e-12, "alpha")
    gpu.assemble_jacobian()
    Bg = gpu.bsr_values()
    Bref = fx["ref_jac_values"].reshape(Bg.shape)
    assert_close_scaled(Bg, Bref, np.abs(Bref).max(axis=(1, 2), keepdims=True), 1e-12, "Jacobian vs reference")


@pytest.mark.parametrize("name", ["plate", "bump"])
def test_mut_matches_oracle_multi_tile(name):
    """Several tiles (the flux kernel's halo side reads mu_t from memory); mu_t
    0..40 laminar viscosities on cells and ghosts."""
    case = small_cases()[name]
    mesh = cfdg.Mesh.generate(**case.mesh)
    gpu = cfdg.Solver(mesh, case.numerics, case.bcs)
    ora = Oracle(mesh, case.numerics, case.bcs)
    U = perturbed_state(ora, case)
    mut = cfdg.cases.sutherland(300.0) * 40.0 * np.random.default_rng(3).random(ora.n_total)
    for s in (gpu, ora):
        s.set_state(U)
        s.set_eddy_viscosity(mut)
        s.eval_residual()
    _check_against(gpu, ora, name)


def test_mut_clear_restores_laminar():
    """Clearing the field gives the laminar residual bit for bit; the module
    and a prescribed field exclude each other."""
    case = small_cases()["plate"]
    mesh = cfdg.Mesh.generate(**case.mesh)
    gpu = cfdg.Solver(mesh, case.numerics, case.bcs)
    U = perturbed_state(gpu, case)
    gpu.set_state(U)
    gpu.eval_residual()
    R0, _ = gpu.residual()
    gpu.set_eddy_viscosity(np.full(gpu.n_total, 1e-4))
    gpu.eval_residual()
    R1, _ = gpu.residual()
    assert not np.array_equal(R0, R1)
    gpu.set_eddy_viscosity(None)
    gpu.eval_residual()
    R2, _ = gpu.residual()
    assert np.array_equal(R0, R2)
    gpu.set_eddy_viscosity(np.full(gpu.n_total, 1e-4))
    gpu.set_turbulence(cfdg.Turbulence.make(case.rho, case.p))
    with pytest.raises(cfdg.CfdgError):
        gpu.set_eddy_viscosity(np.full(gpu.n_total, 1e-4))
