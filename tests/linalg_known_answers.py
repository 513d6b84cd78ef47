"""Known-answer linear systems on the 5x5-block BSR pattern of a mesh, after
the reference's own linear-algebra suite (tests/linalg_test.cpp:404-1510).

The reference builds its matrices on a mesh partition's face graph with block
sizes 1, 3 and 4; the solver's matrices here always have 5x5 blocks, so each
case is restated with 5x5 blocks: a scalar case (bs = 1) becomes the scalar
times the 5x5 identity, a coupled 3x3 block keeps its 3x3 entries in the
block's top-left corner and adds the same diagonal shift to the last two
components.  Rows and columns are mesh cells; "gid" is the mesh cell index,
as in the reference on one rank.  The shared helpers below build the block
values from (row_ptr, cols, diag) and the expected answers; the tests that use
them run on the oracle (test_linalg_known_answers.py) and on the device
(test_gpu_linalg_known_answers.py) through the same calls (set_bsr_values,
spmv, sgs_apply, linear_solve).
"""
import numpy as np

import cfdg
import cfdg.cases as cases

MESHES = {
    "hex": dict(kind=0, ni=9, nj=7, nk=5, x0=0.0, x1=1.0, y0=0.0, y1=1.0, z0=0.0, z1=0.6, jitter=0.15, seed=4,
                reorder=1),
    "tet": dict(kind=1, ni=5, nj=4, nk=3, x0=0.0, x1=2.0, y0=0.0, y1=1.0, z0=0.0, z1=0.8, bump_h=0.05, bump_w=0.3,
                bump_x=1.0, seed=2, reorder=1),
}
I5 = np.eye(5)


def make(factory, name):
    """A solver-like object (oracle or device) on the named mesh; the flow
    numerics only size the context, the tests set every block value."""
    case = cases.ramp(ni=4, nj=2, nk=2)
    mesh = cfdg.Mesh.generate(**MESHES[name])
    bcs = [case.bcs[0]] * mesh.n_patches
    return factory(mesh, case.numerics, bcs)


def blocks(s, diag_block, off_block):
    """Block values in the pattern's order: diag_block(i, degree) for the
    diagonal of row i, off_block(i, j) for column j != i."""
    rp, cols, dg = s.bsr_pattern()
    v = np.zeros((len(cols), 5, 5))
    for i in range(len(rp) - 1):
        deg = rp[i + 1] - rp[i] - 1
        for k in range(rp[i], rp[i + 1]):
            v[k] = diag_block(i, deg) if k == dg[i] else off_block(i, cols[k])
    return v


def apply(s, v, x):
    """A x in numpy from the block values (mesh order), for the expected
    right-hand sides: b = A x* as the reference forms it with mat.apply."""
    rp, cols, _ = s.bsr_pattern()
    X = x.reshape(-1, 5)
    out = np.zeros_like(X)
    for i in range(len(rp) - 1):
        acc = np.zeros(5)
        for k in range(rp[i], rp[i + 1]):
            acc = acc + v[k] @ X[cols[k]]
        out[i] = acc
    return out.ravel()


def flat_gid(n):
    """5 i + v of every vector entry (the per-entry "gid" of the scalar cases)."""
    return np.arange(5 * n, dtype=np.float64)


# --- the matrices ------------------------------------------------------------

def identity(s):
    return blocks(s, lambda i, d: I5, lambda i, j: 0 * I5)


def laplacian(s, shift=0.0):
    """(L + shift I) x I5, L the face-graph Laplacian (linalg_test.cpp:490-518, 1010-1025, 1265-1283)."""
    return blocks(s, lambda i, d: (d + shift) * I5, lambda i, j: -I5)


def adjacency(s):
    """Off-diagonal ones, zero diagonal: A 1 = degree (linalg_test.cpp:520-551)."""
    return blocks(s, lambda i, d: 0 * I5, lambda i, j: I5)


COUPLING = np.array([[2.0, 0.5, 0.0, 0.25, 0.0],
                     [0.5, 3.0, 0.5, 0.0, 0.0],
                     [0.0, 0.5, 2.0, 0.0, 0.5],
                     [0.25, 0.0, 0.0, 1.0, 0.25],
                     [0.0, 0.0, 0.5, 0.25, 4.0]])


def coupled_laplacian(s):
    """degree C on the diagonal, -C off it: A 1 = 0 for any coupling C
    (linalg_test.cpp:594-640, bs = 3 there)."""
    return blocks(s, lambda i, d: d * COUPLING, lambda i, j: -COUPLING)


LU_A = np.array([[4.0, 1.0, 0.5, 0.0, 0.0],
                 [1.0, 5.0, 1.0, 0.0, 0.0],
                 [0.5, 1.0, 3.0, 0.0, 0.0],
                 [0.0, 0.0, 0.0, 2.0, 0.5],
                 [0.0, 0.0, 0.0, 0.25, 6.0]])
LU_X = np.array([1.0, 2.0, 1.5, 0.5, -1.0])
# a block whose factorisation needs row exchanges (the first pivot is not the
# column's largest entry, nor the second)
LU_P = np.array([[0.01, 2.0, 1.0, 0.0, 3.0],
                 [3.0, 0.5, 0.0, 1.0, 0.0],
                 [1.0, 4.0, 0.02, 2.0, 1.0],
                 [0.0, 1.0, 5.0, 0.5, 0.0],
                 [2.0, 0.0, 1.0, 6.0, 0.03]])
LU_PX = np.array([-1.0, 0.5, 2.0, 1.25, -0.75])


def shifted_coupled(s, shift):
    """(degree + shift) I plus the symmetric coupling of linalg_test.cpp:815-823
    in the top-left 3x3, -I off the diagonal."""
    def dblk(i, d):
        b = (d + shift) * I5
        b[0, 1] = b[1, 0] = 0.4
        b[0, 2] = b[2, 0] = 0.2
        b[1, 2] = b[2, 1] = 0.3
        return b
    return blocks(s, dblk, lambda i, j: -I5)


def nonsymmetric(s):
    """-1.6 to later, -0.6 to earlier cells, diagonal = sum |off| + 2.5
    (linalg_test.cpp:1075-1090)."""
    rp, cols, dg = s.bsr_pattern()

    def dblk(i, d):
        off = sum(1.6 if cols[k] > i else 0.6 for k in range(rp[i], rp[i + 1]) if k != dg[i])
        return (off + 2.5) * I5
    return blocks(s, dblk, lambda i, j: (-1.6 if j > i else -0.6) * I5)


def block_nonsymmetric(s):
    """(degree + 4) I plus the non-symmetric coupling of linalg_test.cpp:1146-1150
    in the top-left 3x3; -1.3 I to later, -0.7 I to earlier cells."""
    def dblk(i, d):
        b = (d + 4.0) * I5
        b[0, 1], b[0, 2] = 0.6, -0.4
        b[1, 0], b[1, 2] = -0.5, 0.5
        b[2, 0], b[2, 1] = 0.3, -0.3
        return b
    return blocks(s, dblk, lambda i, j: (-1.3 if j > i else -0.7) * I5)


def tight(precond=cfdg.PRECOND_NONE, order=cfdg.SGS_SEQUENTIAL):
    """The reference's tests run BiCGSTAB with its default rel_tol = 0.1 and
    then ask for a 1e-7 solution error; the tolerance is set explicitly here."""
    lp = cases.default_linparams(rel_tol=1e-12, abs_tol=0.0, max_iter=2000, order=order)
    lp.preconditioner = precond
    return lp


def sgs_iterations(s, b, n, order, x=None):
    """n stationary symmetric Gauss-Seidel iterations x <- x + M^-1 (b - A x)
    (one forward and one backward sweep each, as execute_sgs_iteration,
    linalg_test.cpp:147-173); returns x and the residual norm after each."""
    x = np.zeros_like(b) if x is None else x
    norms = []
    for _ in range(n):
        x = x + s.sgs_apply(order, b - s.spmv(x))
        norms.append(np.linalg.norm(b - s.spmv(x)))
    return x, norms


# --- the checks, shared by the oracle and the device tests ------------------

def check_spmv_identities(s):
    n = s.n_cells
    x = np.cos(0.3 * flat_gid(n)) + 2.0
    s.set_bsr_values(identity(s))
    assert np.array_equal(s.spmv(x), x)
    ones = np.ones(5 * n)
    s.set_bsr_values(laplacian(s))
    assert np.array_equal(s.spmv(ones), np.zeros(5 * n))
    rp = s.bsr_pattern()[0]
    s.set_bsr_values(adjacency(s))
    assert np.array_equal(s.spmv(ones), np.repeat(np.diff(rp) - 1, 5).astype(np.float64))
    s.set_bsr_values(coupled_laplacian(s))
    assert np.array_equal(s.spmv(ones), np.zeros(5 * n))


def check_block_lu(s):
    n = s.n_cells
    for A, X in ((LU_A, LU_X), (LU_P, LU_PX)):
        s.set_bsr_values(blocks(s, lambda i, d: A, lambda i, j: 0 * I5))
        b = np.tile(A @ X, n)
        for order in (cfdg.SGS_SEQUENTIAL, cfdg.SGS_MULTICOLOUR):
            z = s.sgs_apply(order, b).reshape(n, 5)
            err = np.abs(z - X).sum(axis=1)  # per row, as linalg_test.cpp:721-724
            assert err.max() <= 1e-13 * max(1.0, np.abs(X).max()), (order, err.max())


def check_sgs_stationary(s, order):
    n = s.n_cells
    # scalar case: A = L + 2 I, x* = 1, 25 iterations (linalg_test.cpp:730-789)
    v = laplacian(s, 2.0)
    s.set_bsr_values(v)
    xs = np.ones(5 * n)
    b = apply(s, v, xs)
    x, norms = sgs_iterations(s, b, 25, order)
    assert norms[-1] / np.linalg.norm(b) <= 1e-4
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) <= 1e-4
    # coupled case, x* = (1, 2, 3, 4, 5) per cell (linalg_test.cpp:795-866)
    v = shifted_coupled(s, 3.0)
    s.set_bsr_values(v)
    xs = np.tile(np.arange(1.0, 6.0), n)
    b = apply(s, v, xs)
    _, norms = sgs_iterations(s, b, 25, order)
    assert norms[-1] / np.linalg.norm(b) <= 1e-4
    # monotone decrease, b = 1, 20 iterations (linalg_test.cpp:871-935)
    s.set_bsr_values(laplacian(s, 3.0))
    b = np.ones(5 * n)
    _, norms = sgs_iterations(s, b, 20, order)
    prev = np.linalg.norm(b)
    for r in norms:
        assert r <= prev + 1e-12
        prev = r


def check_bicgstab(s, precond, order=cfdg.SGS_SEQUENTIAL):
    n = s.n_cells
    g = flat_gid(n)
    for build, xs in ((lambda: laplacian(s, 3.0), 1.0 + np.cos(0.2 * g)),
                      (lambda: nonsymmetric(s), 1.0 + np.sin(0.1 * g)),
                      (lambda: block_nonsymmetric(s), np.tile(np.arange(1.0, 6.0), n))):
        v = build()
        s.set_bsr_values(v)
        b = apply(s, v, xs)
        x, res = s.linear_solve(tight(precond, order), b=b)
        assert res.status == cfdg.CONVERGED
        assert np.linalg.norm(x - xs) / np.linalg.norm(xs) <= 1e-7
