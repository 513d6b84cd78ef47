"""Failure paths of the linear solve, pinned against the reference.

tests/golden/linsys_identity.npz holds the reference's own
BSRLinearSystem<5>::solve (BiCGSTAB, no preconditioner or SGS) on identity
blocks with b = ones and a zero guess.  With 5 N a perfect square (320
cells) ||b||^2 is exact, alpha = 1, s = t = 0 and the reference returns
Diverged at tt == 0 after one iteration without updating x
(bicgstab.cpp:137-140); with 5 N = 1280 alpha rounds off 1 and the solve
converges in one iteration.  The oracle must reproduce status, iteration
count, norms and solution bit for bit."""
import numpy as np
import pytest

import cfdg
import cfdg.cases as cases
from golden_io import load
from oracle_py import Oracle

KEYS = ["sq_p0", "sq_p1", "nsq_p0", "nsq_p1"]


def linsys_case(fx, key):
    nk = int(fx[f"{key}_mesh_nk"][0])
    case = cases.plate(ni=8, nj=8, nk=nk)
    mesh = cfdg.Mesh.generate(**case.mesh)
    pre, order, rel, ab, mx = fx[f"{key}_lin"]
    lp = cfdg.LinParams()
    lp.preconditioner, lp.sgs_order, lp.rel_tol, lp.abs_tol, lp.max_iter = int(pre), int(order), rel, ab, int(mx)
    return case, mesh, lp


@pytest.mark.parametrize("key", KEYS)
def test_oracle_breakdown_quirk_bit_exact(key):
    fx = load("linsys_identity")
    case, mesh, lp = linsys_case(fx, key)
    o = Oracle(mesh, case.numerics, case.bcs)
    o.set_bsr_values(fx[f"{key}_values"])
    x, res = o.linear_solve(lp, b=fx[f"{key}_b"])
    assert [res.status, res.iterations] == list(fx[f"{key}_ref_lin_status"])
    assert [res.initial_residual, res.final_residual] == list(fx[f"{key}_ref_lin_norms"])
    assert np.array_equal(x, fx[f"{key}_ref_lin_x"])
    if key.startswith("sq"):
        assert res.status == cfdg.DIVERGED and not x.any()
    else:
        assert res.status == cfdg.CONVERGED
