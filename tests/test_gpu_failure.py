"""Failure paths on the device, forced through data only (every kernel exits
normally):

  * BiCGSTAB breakdown at tt == 0: identity blocks, b = ones, 5 N a perfect
    square -> Diverged after one iteration with x unchanged, as the
    reference's own solve (tests/golden/linsys_identity.npz,
    bicgstab.cpp:137-140); 5 N not a square -> converged, x = 1;
  * a non-finite state: NaN written with cfdg_set_state makes cfdg_run_steady
    stop with CFDG_DIVERGED at its first check (solver.hpp:625-629 stops the
    run on a non-finite residual norm);
  * a solve on block storage that a precision change emptied is refused with
    CFDG_ERR_STATE until the blocks are assembled again.
"""
import numpy as np
import pytest

import cfdg
import cfdg.cases as cases
from golden_io import load
from helpers import perturbed_state
from test_failure_paths import KEYS, linsys_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("key", KEYS)
def test_breakdown_quirk_matches_reference(key):
    fx = load("linsys_identity")
    case, mesh, lp = linsys_case(fx, key)
    s = cfdg.Solver(mesh, case.numerics, case.bcs)
    s.set_bsr_values(fx[f"{key}_values"])
    x, res = s.linear_solve(lp, b=fx[f"{key}_b"])
    ref_status, ref_its = fx[f"{key}_ref_lin_status"]
    assert (res.status, res.iterations) == (ref_status, ref_its)
    assert res.initial_residual == fx[f"{key}_ref_lin_norms"][0]
    xr = fx[f"{key}_ref_lin_x"]
    if key.startswith("sq"):
        assert res.status == cfdg.DIVERGED and np.array_equal(x, xr)  # x untouched: all zero
    else:
        assert np.max(np.abs(x - xr)) <= 1e-12


def test_nan_state_stops_steady_run_diverged():
    case = cases.plate(ni=16, nj=8, nk=4)
    mesh = cfdg.Mesh.generate(**case.mesh)
    s = cfdg.Solver(mesh, case.numerics, case.bcs)
    U = perturbed_state(s, case)
    U[0, mesh.n_cells // 2] = np.nan
    s.set_state(U)
    out = s.run_steady(cases.default_linparams(), max_steps=50, rel_tol=1e-8)
    assert out.status == cfdg.DIVERGED
    assert out.steps == 1


def test_solve_refused_after_storage_change():
    case = cases.ramp(ni=12, nj=4, nk=2)
    mesh = cfdg.Mesh.generate(**case.mesh)
    s = cfdg.Solver(mesh, case.numerics, case.bcs)
    s.set_state(perturbed_state(s, case))
    lp = cases.default_linparams()
    s.eval_residual()
    with pytest.raises(cfdg.CfdgError) as e:
        s.linear_solve(lp)  # nothing assembled yet
    assert e.value.code == -3
    s.compute_dt()
    s.assemble_jacobian()
    x64, r64 = s.linear_solve(lp)
    s.set_precision(cfdg.Precision.make(cfdg.PREC_FP32, cfdg.PREC_FP32))
    for call in (lambda: s.linear_solve(lp), lambda: s.spmv(np.ones(5 * mesh.n_cells)),
                 lambda: s.sgs_apply(cfdg.SGS_MULTICOLOUR, np.ones(5 * mesh.n_cells))):
        with pytest.raises(cfdg.CfdgError) as e:
            call()
        assert e.value.code == -3
    s.assemble_jacobian()
    x32, r32 = s.linear_solve(lp)
    assert r32.status == cfdg.CONVERGED
    # same storage, another inner tolerance: the blocks are kept
    s.set_precision(cfdg.Precision.make(cfdg.PREC_FP32, cfdg.PREC_FP32, inner_rel_tol=0.05))
    x32b, r32b = s.linear_solve(lp)
    assert r32b.status == cfdg.CONVERGED
