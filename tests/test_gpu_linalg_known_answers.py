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
