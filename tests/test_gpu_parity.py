"""HIP path (libcfdg.so through the C ABI) against the CPU oracle and against
the reference's own outputs (golden fixtures).

Tolerances (stated per quantity):
  * primitives, gradients, limiters, spectral radius, dt, alpha: 1e-12
    relative (the kernels repeat the oracle's operation order; the only
    non-IEEE-exact operation on the path is pow() in the farfield state);
  * residual: |R_gpu - R_ref| <= 1e-12 * S per cell and variable, S = sum
    over the cell's faces of |F| (the magnitude of the terms summed);
  * Jacobian: per block max|B_gpu - B_ref| <= 1e-12 max|B_ref|;
  * BSR pattern and SGS row groups: identical integers;
  * one sequential-order SGS application and SpMV: 1e-12 relative to max|z|;
  * BiCGSTAB solutions: 1e-9 relative to max|x| (dot products are reduced in
    a different order than the reference's serial loops);
  * states after implicit steps: 1e-9 relative per variable scale.
"""
import numpy as np
import pytest

import cfdg
from golden_io import bcs_of, fixture_names, linparams_of, load, mesh_of, numerics_of
from helpers import assert_close_scaled, perturbed_state, small_cases
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

CASES = small_cases()


def _pair(case):
    mesh = cfdg.Mesh.generate(**case.mesh)
    gpu = cfdg.Solver(mesh, case.numerics, case.bcs)
    ora = Oracle(mesh, case.numerics, case.bcs)
    U = perturbed_state(ora, case)
    gpu.set_state(U)
    ora.set_state(U)
    return mesh, gpu, ora


def _rel_check(a, b, tol, what, floor=0.0):
    scale = np.maximum(np.abs(b), floor)
    d = np.abs(a - b)
    bad = d > tol * scale
    assert not bad.any(), f"{what}: {bad.sum()} entries off (max |d| {d.max():.3e})"


@pytest.mark.parametrize("name", sorted(CASES))
def test_residual_matches_oracle(name):
    case = CASES[name]
    mesh, gpu, ora = _pair(case)
    gpu.eval_residual()
    ora.eval_residual()
    qg, qo = gpu.primitives(), ora.primitives()
    _rel_check(qg, qo, 1e-12, "primitives")
    gg, pg = gpu.gradients()
    go, po = ora.gradients()
    gscale = np.abs(go).max(axis=1, keepdims=True)
    assert_close_scaled(gg, go, gscale, 1e-12, "gradients")
    _rel_check(pg, po, 1e-12, "limiters", floor=1e-300)
    Rg, lg = gpu.residual()
    Ro, lo = ora.residual()
    S = ora.flux_scale()
    assert_close_scaled(Rg, Ro, S, 1e-12, "residual")
    _rel_check(lg, lo, 1e-12, "spectral radius")
    print(f"{name}: bit-identical residual entries {np.mean(Rg == Ro):.4f}")


@pytest.mark.parametrize("name", sorted(CASES))
def test_jacobian_matches_oracle(name):
    case = CASES[name]
    mesh, gpu, ora = _pair(case)
    for s in (gpu, ora):
        s.eval_residual()
        s.compute_dt()
        s.assemble_jacobian()
    dtg, ag = gpu.dt()
    dto, ao = ora.dt()
    _rel_check(dtg, dto, 1e-12, "dt")
    _rel_check(ag, ao, 1e-12, "alpha")
    for a, b in zip(gpu.bsr_pattern(), ora.bsr_pattern()):
        assert np.array_equal(a, b)
    Bg, Bo = gpu.bsr_values(), ora.bsr_values()
    bmax = np.abs(Bo).max(axis=(1, 2), keepdims=True)
    assert_close_scaled(Bg, Bo, bmax, 1e-12, "Jacobian blocks")


@pytest.mark.parametrize("name", sorted(CASES))
def test_sgs_and_spmv_match_oracle(name):
    case = CASES[name]
    mesh, gpu, ora = _pair(case)
    for s in (gpu, ora):
        s.eval_residual()
        s.compute_dt()
        s.assemble_jacobian()
    rng = np.random.default_rng(5)
    r = rng.standard_normal(5 * mesh.n_cells)
    for order in (cfdg.SGS_SEQUENTIAL, cfdg.SGS_MULTICOLOUR):
        zg = gpu.sgs_apply(order, r)
        zo = ora.sgs_apply(order, r)
        assert_close_scaled(zg, zo, np.abs(zo).max(), 1e-12, f"SGS order {order}")
    yg, yo = gpu.spmv(r), ora.spmv(r)
    assert_close_scaled(yg, yo, np.abs(yo).max(), 1e-12, "SpMV")


@pytest.mark.parametrize("name", sorted(CASES))
def test_sgs_row_groups_match_oracle(name):
    """The rows each device sweep is handed, read back from the device: the
    forward multicolour groups are the oracle's colours, the backward ones
    the colours reversed; the sequential groups are the dependency levels of
    the natural forward / backward sweeps.  Identical integers."""
    case = CASES[name]
    mesh, gpu, ora = _pair(case)
    col, ncol, fl, nf, bl, nb = ora.row_groups()
    f, b, n = gpu.sgs_row_groups(cfdg.SGS_MULTICOLOUR)
    assert n == (ncol, ncol)
    assert np.array_equal(f, col)
    assert np.array_equal(b, ncol - 1 - col)
    f, b, n = gpu.sgs_row_groups(cfdg.SGS_SEQUENTIAL)
    assert n == (nf, nb)
    assert np.array_equal(f, fl)
    assert np.array_equal(b, bl)


@pytest.mark.parametrize("name", sorted(CASES))
def test_multicolour_sgs_matches_numpy_colour_order(name):
    """Device multicolour SGS against an independent numpy block-SGS run in
    colour order with no skipped group (tests/test_sgs_oracle.py pins that
    numpy sweep to the reference-pinned sequential one): 1e-12 of max|z|."""
    from test_sgs_oracle import numpy_sgs
    case = CASES[name]
    mesh, gpu, ora = _pair(case)
    for s in (gpu, ora):
        s.eval_residual()
        s.compute_dt()
        s.assemble_jacobian()
    rp, cols, dg = ora.bsr_pattern()
    B = ora.bsr_values()
    col = ora.row_groups()[0]
    rows = np.arange(mesh.n_cells)
    r = np.random.default_rng(6).standard_normal(5 * mesh.n_cells)
    zn = numpy_sgs(rp, cols, dg, B, r, rows[np.lexsort((rows, col))], rows[np.lexsort((rows, -col))])
    zg = gpu.sgs_apply(cfdg.SGS_MULTICOLOUR, r)
    assert_close_scaled(zg, zn, np.abs(zn).max(), 1e-12, "multicolour SGS vs numpy")


@pytest.mark.parametrize("order", [cfdg.SGS_SEQUENTIAL, cfdg.SGS_MULTICOLOUR])
@pytest.mark.parametrize("name", sorted(CASES))
def test_linear_solve_matches_oracle(name, order):
    case = CASES[name]
    mesh, gpu, ora = _pair(case)
    for s in (gpu, ora):
        s.eval_residual()
        s.compute_dt()
        s.assemble_jacobian()
    lp = cfdg.cases.default_linparams(rel_tol=1e-6, max_iter=200, order=order)
    xg, rg = gpu.linear_solve(lp)
    xo, ro = ora.linear_solve(lp)
    assert rg.status == ro.status == cfdg.CONVERGED
    assert abs(rg.iterations - ro.iterations) <= 1
    assert_close_scaled(xg, xo, np.abs(xo).max(), 1e-6 if rg.iterations != ro.iterations else 1e-9, "solution")


@pytest.mark.parametrize("name", sorted(CASES))
def test_implicit_steps_match_oracle(name):
    case = CASES[name]
    mesh = cfdg.Mesh.generate(**case.mesh)
    gpu = cfdg.Solver(mesh, case.numerics, case.bcs)
    ora = Oracle(mesh, case.numerics, case.bcs)
    for s in (gpu, ora):
        s.init_uniform(case.rho, case.velocity, case.p)
    lp = cfdg.cases.default_linparams(order=cfdg.SGS_SEQUENTIAL)
    for it in range(5):
        sg = gpu.implicit_step(lp)
        so = ora.implicit_step(lp)
        assert np.allclose(sg.l2, so.l2, rtol=1e-9, atol=0.0), (it, list(sg.l2), list(so.l2))
    Ug, Uo = gpu.get_state(), ora.get_state()
    scale = np.abs(Uo).max(axis=1, keepdims=True)
    assert_close_scaled(Ug, Uo, scale, 1e-9, "state after 5 implicit steps")


@pytest.mark.parametrize("name", fixture_names())
def test_against_reference_golden(name):
    """The HIP path against the reference's own outputs."""
    fx = load(name)
    mesh = mesh_of(fx)
    gpu = cfdg.Solver(mesh, numerics_of(fx), bcs_of(fx))
    gpu.set_state(fx["U"])
    gpu.eval_residual()
    q = gpu.primitives().ravel()
    _rel_check(q, fx["ref_q"], 1e-12, "primitives")
    g, phi = gpu.gradients()
    go = fx["ref_grad"].reshape(g.shape)
    assert_close_scaled(g, go, np.abs(go).max(axis=1, keepdims=True), 1e-12, "gradients")
    _rel_check(phi.ravel(), fx["ref_phi"], 1e-12, "limiters")
    R, lam = gpu.residual()
    Ro = fx["ref_res"].reshape(R.shape)
    assert_close_scaled(R, Ro, np.abs(Ro).max(axis=1, keepdims=True), 1e-12, "residual")
    gpu.compute_dt()
    gpu.assemble_jacobian()
    rp, cols, dg = gpu.bsr_pattern()
    assert np.array_equal(rp, fx["ref_row_ptr"]) and np.array_equal(cols, fx["ref_cols"])
    assert np.array_equal(dg, fx["ref_diag"])
    B = gpu.bsr_values()
    Bo = fx["ref_values"].reshape(B.shape)
    assert_close_scaled(B, Bo, np.abs(Bo).max(axis=(1, 2), keepdims=True), 1e-12, "Jacobian")
    du, res = gpu.linear_solve(linparams_of(fx))
    assert res.status == int(fx["ref_lin_status"][0])
    assert abs(res.iterations - int(fx["ref_lin_status"][1])) <= 1
    assert_close_scaled(du, fx["ref_du"], np.abs(fx["ref_du"]).max(), 1e-6, "linear solution")


@pytest.mark.parametrize("name", sorted(CASES))
def test_residual_is_deterministic(name):
    case = CASES[name]
    mesh, gpu, _ = _pair(case)
    gpu.eval_residual()
    R1, l1 = gpu.residual()
    gpu.eval_residual()
    R2, l2 = gpu.residual()
    assert np.array_equal(R1, R2) and np.array_equal(l1, l2)


@pytest.mark.parametrize("name", sorted(CASES))
def test_boundary_integrals_match_oracle(name):
    """Per-patch mass, energy and force: |gpu - oracle| <= 1e-12 x (sum over
    all boundary faces of |F|) -- the device sums each patch in a fixed tree
    order instead of the reference's face order."""
    case = CASES[name]
    mesh, gpu, ora = _pair(case)
    gpu.eval_residual()
    ora.eval_residual()
    mg, eg, fg = gpu.boundary_integrals()
    mo, eo, fo = ora.boundary_integrals()
    S = ora.flux_scale().sum(axis=1)
    assert np.all(np.abs(mg - mo) <= 1e-12 * S[0] + 1e-300)
    assert np.all(np.abs(eg - eo) <= 1e-12 * S[4] + 1e-300)
    for k in range(3):
        assert np.all(np.abs(fg[:, k] - fo[:, k]) <= 1e-12 * S[1 + k] + 1e-300)


@pytest.mark.parametrize("name", fixture_names())
def test_boundary_integrals_match_reference_golden(name):
    fx = load(name)
    mesh = mesh_of(fx)
    gpu = cfdg.Solver(mesh, numerics_of(fx), bcs_of(fx))
    gpu.set_state(fx["U"])
    gpu.eval_residual()
    mg, eg, _ = gpu.boundary_integrals()
    scale = np.abs(fx["ref_res"]).reshape(5, -1).sum(axis=1)
    assert np.all(np.abs(mg - fx["ref_bnd_mass"]) <= 1e-12 * scale[0] + 1e-300)
    assert np.all(np.abs(eg - fx["ref_bnd_energy"]) <= 1e-12 * scale[4] + 1e-300)


@pytest.mark.parametrize("max_iter", [0, 1, 2, 5])
def test_linear_solve_iteration_limit_matches_oracle(max_iter):
    """The device-driven BiCGSTAB stops where the reference's loop stops
    (bicgstab.cpp:97-201): the iteration limit (status MAX_ITERATIONS after
    exactly max_iter iterations), the same iterate (1e-9) and residual."""
    case = CASES["plate"]
    mesh, gpu, ora = _pair(case)
    for s in (gpu, ora):
        s.eval_residual()
        s.compute_dt()
        s.assemble_jacobian()
    lp = cfdg.cases.default_linparams(rel_tol=1e-14, abs_tol=0.0, max_iter=max_iter, order=cfdg.SGS_SEQUENTIAL)
    xg, rg = gpu.linear_solve(lp)
    xo, ro = ora.linear_solve(lp)
    assert rg.status == ro.status == cfdg.MAX_ITERATIONS
    assert rg.iterations == ro.iterations == max_iter
    if max_iter:
        assert_close_scaled(xg, xo, np.abs(xo).max(), 1e-9, "iterate at the limit")
    else:
        assert not np.any(xg)
    assert abs(rg.final_residual - ro.final_residual) <= 1e-9 * ro.initial_residual


def test_linear_solve_zero_rhs_and_tolerance_edges():
    """b = 0 converges before the first iteration (tolerance max(abs, rel ||b||)
    with ||b|| = 0 -> the scale 1 of bicgstab.cpp:69-72); a loose tolerance met
    by the initial residual also stops with zero iterations."""
    case = CASES["plate"]
    mesh, gpu, ora = _pair(case)
    for s in (gpu, ora):
        s.eval_residual()
        s.compute_dt()
        s.assemble_jacobian()
    n = 5 * gpu.n_cells
    lp = cfdg.cases.default_linparams(order=cfdg.SGS_SEQUENTIAL)
    xg, rg = gpu.linear_solve(lp, b=np.zeros(n))
    xo, ro = ora.linear_solve(lp, b=np.zeros(n))
    assert rg.status == ro.status == cfdg.CONVERGED and rg.iterations == ro.iterations == 0
    assert not np.any(xg)
    loose = cfdg.cases.default_linparams(rel_tol=2.0, order=cfdg.SGS_SEQUENTIAL)
    xg, rg = gpu.linear_solve(loose)
    xo, ro = ora.linear_solve(loose)
    assert rg.status == ro.status == cfdg.CONVERGED and rg.iterations == ro.iterations == 0


def test_singular_diagonal_block_is_reported():
    """A zero diagonal block makes the SGS set-up fail (preconditioners.cpp:84-89
    is fatal in the reference): the solve reports CFDG_ERR_SINGULAR naming the
    mesh row, for the solve and for a single preconditioner application."""
    case = CASES["plate"]
    mesh, gpu, _ = _pair(case)
    gpu.eval_residual()
    gpu.compute_dt()
    gpu.assemble_jacobian()
    rp, cols, dg = gpu.bsr_pattern()
    vals = gpu.bsr_values()
    row = 7
    vals[dg[row]] = 0.0
    gpu.set_bsr_values(vals)
    lp = cfdg.cases.default_linparams(order=cfdg.SGS_SEQUENTIAL)
    with pytest.raises(cfdg.CfdgError) as e:
        gpu.linear_solve(lp, b=np.ones(5 * gpu.n_cells))
    assert e.value.code == cfdg.ERR_SINGULAR and f"row {row}" in str(e.value)
    with pytest.raises(cfdg.CfdgError) as e:
        gpu.sgs_apply(cfdg.SGS_SEQUENTIAL, np.ones(5 * gpu.n_cells))
    assert e.value.code == cfdg.ERR_SINGULAR
