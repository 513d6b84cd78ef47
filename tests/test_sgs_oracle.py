"""Multicolour SGS pinned against the reference-pinned sequential sweep.

The oracle's sequential SGS reproduces the reference's bsr_sgs_sweep
(bsr_helpers.hpp:390-435) bit for bit (through the golden BiCGSTAB+SGS
solutions).  Multicolour SGS is the same symmetric Gauss-Seidel sweep in
another row order.  Here an independent numpy block-SGS that takes any row
order (np.linalg.solve for the diagonal blocks, no skipped group) is first
checked against the oracle's sequential sweep in the natural order, and then
in colour order against the oracle's multicolour apply, which skips the
backward sweep's first group.  The colouring and the level sets are checked
for validity on the face graph.  Tolerance: 1e-12 of max|z| (LAPACK's LU and
the oracle's block LU round differently)."""
import numpy as np
import pytest

import cfdg
from helpers import perturbed_state, small_cases
from oracle_py import Oracle

CASES = small_cases()


def _assembled(name):
    case = CASES[name]
    mesh = cfdg.Mesh.generate(**case.mesh)
    o = Oracle(mesh, case.numerics, case.bcs)
    o.set_state(perturbed_state(o, case))
    o.eval_residual()
    o.compute_dt()
    o.assemble_jacobian()
    return mesh, o


def numpy_sgs(row_ptr, cols, diag, B, r, fwd_order, bwd_order):
    """z = 0; z_i = D_i^-1 (r_i - sum_{k != i} A_ik z_k) over fwd_order, then bwd_order."""
    n = len(diag)
    z = np.zeros((n, 5))
    R = r.reshape(n, 5)
    for order in (fwd_order, bwd_order):
        for i in order:
            s = R[i].copy()
            for k in range(row_ptr[i], row_ptr[i + 1]):
                if k != diag[i]:
                    s -= B[k] @ z[cols[k]]
            z[i] = np.linalg.solve(B[diag[i]], s)
    return z.ravel()


def _face_edges(mesh):
    o, n = mesh.array("face_owner"), mesh.array("face_neigh")
    inner = n >= 0
    return o[inner], n[inner]


@pytest.mark.parametrize("name", sorted(CASES))
def test_colouring_and_levels_valid(name):
    mesh, o = _assembled(name)
    col, ncol, fl, nf, bl, nb = o.row_groups()
    a, b = _face_edges(mesh)
    assert np.all(col[a] != col[b]), "face neighbours share a colour"
    assert col.min() == 0 and col.max() == ncol - 1
    if name != "bump":
        assert ncol == 2  # hexahedral face graph is bipartite
    # forward level of a row exceeds that of every lower neighbour; backward of every higher one
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    assert np.all(fl[hi] > fl[lo])
    assert np.all(bl[lo] > bl[hi])
    assert fl.max() == nf - 1 and bl.max() == nb - 1


@pytest.mark.parametrize("name", sorted(CASES))
def test_sequential_sgs_equals_numpy_natural_order(name):
    mesh, o = _assembled(name)
    rp, cols, dg = o.bsr_pattern()
    B = o.bsr_values()
    r = np.random.default_rng(5).standard_normal(5 * mesh.n_cells)
    n = mesh.n_cells
    zn = numpy_sgs(rp, cols, dg, B, r, range(n), range(n - 1, -1, -1))
    zo = o.sgs_apply(cfdg.SGS_SEQUENTIAL, r)
    assert np.max(np.abs(zo - zn)) <= 1e-12 * np.abs(zn).max()


@pytest.mark.parametrize("name", sorted(CASES))
def test_multicolour_sgs_equals_numpy_colour_order(name):
    mesh, o = _assembled(name)
    rp, cols, dg = o.bsr_pattern()
    B = o.bsr_values()
    col = o.row_groups()[0]
    n = mesh.n_cells
    rows = np.arange(n)
    fwd = rows[np.lexsort((rows, col))]     # colour 0 first, ascending rows within a colour
    bwd = rows[np.lexsort((rows, -col))]    # last colour first
    r = np.random.default_rng(6).standard_normal(5 * n)
    zn = numpy_sgs(rp, cols, dg, B, r, fwd, bwd)
    zo = o.sgs_apply(cfdg.SGS_MULTICOLOUR, r)
    assert np.max(np.abs(zo - zn)) <= 1e-12 * np.abs(zn).max()
    # planted mistake: a backward sweep that visits the colours in forward order
    zw = numpy_sgs(rp, cols, dg, B, r, fwd, fwd)
    assert np.max(np.abs(zw - zn)) > 1e-6 * np.abs(zn).max()
