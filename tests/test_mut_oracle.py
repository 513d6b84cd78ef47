"""The eddy-viscosity branch of the mean viscous flux and Jacobian, pinned
against the reference's own compiled code (tests/golden/plate_mut.npz, written
by tests/golden/make_golden.py mut).

The fixture holds a viscous plate with a seeded, non-uniform mu_t field on
every cell and BC ghost (0..40 laminar viscosities).  The reference ran
ResidualKernel::apply and boundary_integrals with that field as their mut
argument (residual_kernel.hpp:64-169) and JacobianKernel::apply with it and
a given diagonal shift (jacobian_kernel.hpp:74-270, mu_t = 0 on the ghost
side).  The oracle must reproduce every array bit for bit.  The planted
mistakes below (in the oracle, on the CPU) show that the fixture can tell a
wrong turbulent Prandtl number or a dropped mu_t apart."""
import numpy as np

from golden_io import bcs_of, load, mesh_of, numerics_of
from oracle_py import Oracle


def _oracle(fx, numerics=None, mut=True):
    o = Oracle(mesh_of(fx), numerics or numerics_of(fx), bcs_of(fx))
    o.set_state(fx["U"])
    if mut:
        o.set_eddy_viscosity(fx["mut"])
    o.eval_residual()
    return o


def test_mut_residual_bit_exact():
    fx = load("plate_mut")
    o = _oracle(fx)
    q = o.primitives()
    g, phi = o.gradients()
    R, lam = o.residual()
    assert np.array_equal(q.ravel(), fx["ref_q"])
    assert np.array_equal(g.ravel(), fx["ref_grad"])
    assert np.array_equal(phi.ravel(), fx["ref_phi"])
    assert np.array_equal(R.ravel(), fx["ref_res"])
    assert np.array_equal(lam, fx["ref_lam"])
    mass, energy, _ = o.boundary_integrals()
    assert np.array_equal(mass, fx["ref_bnd_mass"])
    assert np.array_equal(energy, fx["ref_bnd_energy"])


def test_mut_jacobian_bit_exact():
    fx = load("plate_mut")
    o = _oracle(fx)
    o.compute_dt()
    _, alpha = o.dt()
    assert np.array_equal(alpha, fx["jac_alpha"])
    o.assemble_jacobian()
    assert np.array_equal(o.bsr_values().ravel(), fx["ref_jac_values"])


def test_mut_fixture_discriminates():
    """Planted mistakes: no mu_t at all, and Pr_t = 0.85 instead of 0.9, each
    move the residual, the spectral radius or the Jacobian well past round-off."""
    fx = load("plate_mut")
    ref_R = fx["ref_res"].reshape(5, -1)
    o = _oracle(fx, mut=False)
    R, lam = o.residual()
    S = o.flux_scale()
    assert np.max(np.abs(R - ref_R) / S) > 1e-3
    assert np.max(np.abs(lam - fx["ref_lam"]) / fx["ref_lam"]) > 1e-3
    num = numerics_of(fx)
    num.prandtl_turb = 0.85
    o = _oracle(fx, numerics=num)
    R, _ = o.residual()
    # only the energy equation carries Pr_t
    assert np.array_equal(R[:4], ref_R[:4])
    assert np.max(np.abs(R[4] - ref_R[4]) / S[4]) > 1e-6
    o.compute_dt()
    o.assemble_jacobian()
    B = o.bsr_values().ravel()
    assert not np.array_equal(B, fx["ref_jac_values"])


def test_ghost_mut_reaches_residual_not_jacobian():
    """The residual reads mu_t of the BC ghost (residual_kernel.hpp:153-154,
    284-285); the Jacobian takes zero there (jacobian_kernel.hpp:252-254).
    Zeroing the ghost values changes the first and not the second."""
    fx = load("plate_mut")
    mut = fx["mut"].copy()
    nc = fx["U"].shape[1]
    mut[nc:] = 0.0
    o = Oracle(mesh_of(fx), numerics_of(fx), bcs_of(fx))
    o.set_state(fx["U"])
    o.set_eddy_viscosity(mut)
    o.eval_residual()
    R, _ = o.residual()
    assert not np.array_equal(R.ravel(), fx["ref_res"])
    # same Jacobian blocks with the fixture's diagonal shift: set it via the values
    o.compute_dt()
    o.assemble_jacobian()
    B = o.bsr_values()
    ref = fx["ref_jac_values"].reshape(B.shape)
    # diagonal blocks differ only through alpha (spectral radius changed), off-diagonal ones are equal
    rp, cols, dg = o.bsr_pattern()
    off = np.ones(B.shape[0], bool)
    off[dg] = False
    assert np.array_equal(B[off], ref[off])
