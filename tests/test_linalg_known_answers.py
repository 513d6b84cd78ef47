"""The oracle's linear algebra against the known answers of the reference's
own linear-algebra suite (tests/linalg_test.cpp:404-1510), restated for 5x5
blocks in tests/linalg_known_answers.py: SpMV identities (identity, Laplacian
null space, row degrees, coupled Laplacian null space), the diagonal-block LU
(the reference's 3x3 case and a block that needs row exchanges), stationary
SGS iterations (25 iterations to a 1e-4 reduction, monotone decrease) and
BiCGSTAB to a 1e-7 solution error on a shifted Laplacian, a non-symmetric and
a block non-symmetric system, unpreconditioned and with SGS.  Runs on the CPU."""
import numpy as np
import pytest

import cfdg
import linalg_known_answers as ka
from oracle_py import Oracle


@pytest.fixture(scope="module", params=sorted(ka.MESHES))
def ora(request):
    return ka.make(Oracle, request.param)


def test_spmv_identities(ora):
    ka.check_spmv_identities(ora)


def test_block_lu(ora):
    ka.check_block_lu(ora)


@pytest.mark.parametrize("order", [cfdg.SGS_SEQUENTIAL, cfdg.SGS_MULTICOLOUR])
def test_sgs_stationary_iterations(ora, order):
    ka.check_sgs_stationary(ora, order)


def test_bicgstab_unpreconditioned(ora):
    ka.check_bicgstab(ora, cfdg.PRECOND_NONE)


@pytest.mark.parametrize("order", [cfdg.SGS_SEQUENTIAL, cfdg.SGS_MULTICOLOUR])
def test_bicgstab_sgs(ora, order):
    ka.check_bicgstab(ora, cfdg.PRECOND_SGS, order)


def test_a_wrong_answer_fails(ora):
    """The block check is sensitive: factors of a block that differs from the
    one the right-hand side was formed with in one off-diagonal entry miss
    the known answer by far more than the 1e-13 bound."""
    wrong = ka.LU_A.copy()
    wrong[0, 1] += 0.5
    ora.set_bsr_values(ka.blocks(ora, lambda i, d: wrong, lambda i, j: 0 * ka.I5))
    b = np.tile(ka.LU_A @ ka.LU_X, ora.n_cells)
    z = ora.sgs_apply(cfdg.SGS_SEQUENTIAL, b).reshape(ora.n_cells, 5)
    assert np.abs(z - ka.LU_X).sum(axis=1).max() > 1e-3
