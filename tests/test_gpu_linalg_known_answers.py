"""The device's linear algebra (cfdg_set_bsr_values, cfdg_spmv,
cfdg_sgs_apply, cfdg_linear_solve) against the known answers of the
reference's own linear-algebra suite (tests/linalg_test.cpp:404-1510),
restated for 5x5 blocks in tests/linalg_known_answers.py; the same checks run
on the oracle in tests/test_linalg_known_answers.py."""
import pytest

import cfdg
import linalg_known_answers as ka

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=sorted(ka.MESHES))
def dev(request):
    print(f"linalg known answers on the {request.param} mesh", flush=True)
    return ka.make(cfdg.Solver, request.param)


def test_spmv_identities(dev):
    ka.check_spmv_identities(dev)


def test_block_lu(dev):
    ka.check_block_lu(dev)


@pytest.mark.parametrize("order", [cfdg.SGS_SEQUENTIAL, cfdg.SGS_MULTICOLOUR])
def test_sgs_stationary_iterations(dev, order):
    ka.check_sgs_stationary(dev, order)


def test_bicgstab_unpreconditioned(dev):
    ka.check_bicgstab(dev, cfdg.PRECOND_NONE)


@pytest.mark.parametrize("order", [cfdg.SGS_SEQUENTIAL, cfdg.SGS_MULTICOLOUR])
def test_bicgstab_sgs(dev, order):
    ka.check_bicgstab(dev, cfdg.PRECOND_SGS, order)


@pytest.mark.parametrize("max_iter", [2000, 31, 32, 33])
def test_long_solves_match_oracle(max_iter):
    """A solve of more than 32 iterations (the device recomputes the true
    residual norm every 32nd iteration, and iterations are enqueued in
    batches sized by the previous solve), and solves stopped at the iteration
    limit around that point: the device gives the oracle's status and
    iteration count and its iterate (1e-9).  Shifted Laplacian (shift 0.001)
    on the tetrahedral mesh with multicolour SGS: 35 iterations to 1e-12."""
    from oracle_py import Oracle
    import numpy as np

    dev = ka.make(cfdg.Solver, "tet")
    ora = ka.make(Oracle, "tet")
    v = ka.laplacian(ora, 0.001)
    xs = 1.0 + np.cos(0.2 * ka.flat_gid(ora.n_cells))
    b = ka.apply(ora, v, xs)
    lp = ka.tight(cfdg.PRECOND_SGS, cfdg.SGS_MULTICOLOUR)
    lp.max_iter = max_iter
    out = []
    for s in (dev, ora):
        s.set_bsr_values(v)
        for _ in range(2):  # the second solve starts with a batch sized by the first
            x, r = s.linear_solve(lp, b=b)
        out.append((x, r))
    (xg, rg), (xo, ro) = out
    assert rg.status == ro.status
    if max_iter < 2000:
        assert rg.status == cfdg.MAX_ITERATIONS and rg.iterations == ro.iterations == max_iter
    else:
        assert rg.status == cfdg.CONVERGED and ro.iterations > 32 and abs(rg.iterations - ro.iterations) <= 1
        assert np.linalg.norm(xg - xs) / np.linalg.norm(xs) <= 1e-7
    assert np.linalg.norm(xg - xo) <= 1e-9 * np.linalg.norm(xo)
