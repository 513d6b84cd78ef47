"""The HIP path at the benchmark's sizes and on the kernel forms that only
large or tetrahedral meshes reach, against the oracle.

* The 2.1M-cell laminar plate of the benchmark (256x128x64 hexahedra, 8,192
  tiles spread over the eight XCD ranges): residual |dR| <= 1e-12 of the
  per-cell flux scale, primitives and limiters 1e-12 relative, gradients
  1e-12 of each component's maximum; Jacobian blocks max|dB| <= 1e-12 of the
  block's maximum.
* Tetrahedral bumps whose node stencils take each gradient kernel form: a
  mesh whose tiles fit the wide LDS form whole, one whose tiles must be split
  into sub-ranges, the hexahedral plate in the 640-slot form, and
  Green-Gauss gradients in the gather form; the form is asserted through
  cfdg_gradient_form, then gradients, limiters and residual are checked as
  above.
* The supersonic ramp run for the 200 implicit steps of its configuration
  (CFL 10, BiCGSTAB + the reference's sequential SGS): state within 1e-8 of
  each variable's scale and residual norms within 1e-6 relative at every
  step (each BiCGSTAB dot product is a tree reduction, 1e-9 per solve).
* The transonic SA bump at reduced size in the forms the benchmark times
  (FP32 Krylov, BF16 blocks, FP32 Krylov with a warm start): every step's
  linear solve converges, the relative residual after 20 steps is within a
  factor 2 of the FP64 run's, and the warm start is judged on every solve.
"""
import numpy as np
import pytest

import cfdg
import cfdg.cases as cases
from helpers import assert_close_scaled, perturbed_state
from oracle_py import Oracle

pytestmark = pytest.mark.gpu


def _rel_check(a, b, tol, what, floor=0.0):
    scale = np.maximum(np.abs(b), floor)
    d = np.abs(a - b)
    bad = d > tol * scale
    assert not bad.any(), f"{what}: {bad.sum()} entries off (max |d| {d.max():.3e})"


def _pair(case, seed=1):
    mesh = cfdg.Mesh.generate(**case.mesh)
    gpu = cfdg.Solver(mesh, case.numerics, case.bcs)
    ora = Oracle(mesh, case.numerics, case.bcs)
    U = perturbed_state(ora, case, seed=seed)
    gpu.set_state(U)
    ora.set_state(U)
    return mesh, gpu, ora


def _check_residual(gpu, ora, what):
    gpu.eval_residual()
    ora.eval_residual()
    _rel_check(gpu.primitives(), ora.primitives(), 1e-12, f"{what} primitives")
    gg, pg = gpu.gradients()
    go, po = ora.gradients()
    assert_close_scaled(gg, go, np.abs(go).max(axis=1, keepdims=True), 1e-12, f"{what} gradients")
    _rel_check(pg, po, 1e-12, f"{what} limiters", floor=1e-300)
    Rg, lg = gpu.residual()
    Ro, lo = ora.residual()
    assert_close_scaled(Rg, Ro, ora.flux_scale(), 1e-12, f"{what} residual")
    _rel_check(lg, lo, 1e-12, f"{what} spectral radius")


def test_full_size_plate_residual_and_jacobian():
    case = cases.plate()  # 256 x 128 x 64, the benchmark's mesh
    mesh, gpu, ora = _pair(case)
    assert mesh.n_cells == 2097152
    form, ngt, nft = gpu.gradient_form()
    assert (form, ngt, nft) == (cfdg.GRAD_LDS, 8192, 8192)
    _check_residual(gpu, ora, "2.1M plate")
    for s in (gpu, ora):
        s.compute_dt()
        s.assemble_jacobian()
    Bg = gpu.bsr_values()
    Bo = ora.bsr_values()
    bmax = np.abs(Bo).max(axis=(1, 2), keepdims=True)
    assert_close_scaled(Bg, Bo, bmax, 1e-12, "2.1M plate Jacobian blocks")


@pytest.mark.parametrize("geometry, form, split", [
    ("bump12_lsq", cfdg.GRAD_LDS_WIDE, False),
    ("bump30_lsq", cfdg.GRAD_LDS_WIDE, True),
    ("plate_lsq", cfdg.GRAD_LDS, False),
    ("bump12_gg", cfdg.GRAD_GATHER, False),
])
def test_gradient_kernel_forms(geometry, form, split):
    if geometry == "bump12_lsq":
        case = cases.bump(ni=12, nj=12, nk=12)
    elif geometry == "bump30_lsq":
        case = cases.bump(ni=30, nj=30, nk=30)
    elif geometry == "plate_lsq":
        case = cases.plate(ni=48, nj=32, nk=16)
    else:
        case = cases.bump(ni=12, nj=12, nk=12, gradient="GREEN_GAUSS_FACE")
    mesh, gpu, ora = _pair(case)
    f, ngt, nft = gpu.gradient_form()
    assert f == form, (geometry, f)
    if form != cfdg.GRAD_GATHER:
        assert (ngt > nft) == split, (geometry, ngt, nft)
    _check_residual(gpu, ora, geometry)


def test_ramp_200_implicit_steps():
    case = cases.ramp()
    mesh = cfdg.Mesh.generate(**case.mesh)
    gpu = cfdg.Solver(mesh, case.numerics, case.bcs)
    ora = Oracle(mesh, case.numerics, case.bcs)
    for s in (gpu, ora):
        s.init_uniform(case.rho, case.velocity, case.p)
    lp = cases.default_linparams(order=cfdg.SGS_SEQUENTIAL)
    for k in range(case.info["steps"]):
        lg = np.array(gpu.implicit_step(lp).l2)
        lo = np.array(ora.implicit_step(lp).l2)
        np.testing.assert_allclose(lg, lo, rtol=1e-6, atol=1e-12 * lo.max(), err_msg=f"norms at step {k + 1}")
    Ug, Uo = gpu.get_state(), ora.get_state()
    scale = np.abs(Uo).max(axis=1, keepdims=True)
    assert_close_scaled(Ug, Uo, scale, 1e-8, "ramp state after 200 steps")


def test_sa_bump_forms_reduced_size():
    case = cases.bump_sa(ni=10, nj=10, nk=10)
    mesh = cfdg.Mesh.generate(**case.mesh)
    s = cfdg.Solver(mesh, case.numerics, case.bcs)
    U0 = perturbed_state(s, case, seed=1234)
    s.set_turbulence(cases.turbulence(case, update=cfdg.TURB_FROZEN))
    rnu0 = s.turbulence()["rho_nu"]
    lp = cases.default_linparams()
    fp32 = cfdg.Precision.make(cfdg.PREC_FP32, cfdg.PREC_FP32, max_outer=8, inner_rel_tol=0.1)
    bf16 = cfdg.Precision.make(cfdg.PREC_BF16, cfdg.PREC_FP32, max_outer=8, inner_rel_tol=0.1)
    net = cfdg.Gnn(seed=3)
    rel = {}
    for name, prec, warm in (("fp64", cfdg.Precision.make(), None), ("fp32", fp32, None), ("bf16", bf16, None),
                             ("fp32_learned", fp32, net)):
        s.set_precision(prec)
        s.set_state(U0)
        s.set_turbulence_state(rnu0)
        s.set_warm_start(warm)
        statuses = []
        l0 = None
        for _ in range(20):
            st = s.implicit_step(lp)
            statuses.append(st.linear.status)
            l2 = np.array(st.l2)
            l0 = l2 if l0 is None else l0
        rel[name] = float(np.max(l2 / l0))
        assert all(x == cfdg.CONVERGED for x in statuses), (name, statuses)
        if warm is not None:
            kept, rejected = s.warm_start_stats()
            assert kept + rejected == 20
    s.set_warm_start(None)
    for name in ("fp32", "bf16", "fp32_learned"):
        assert rel[name] <= 2.0 * rel["fp64"], rel
