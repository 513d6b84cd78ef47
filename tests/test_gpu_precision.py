"""Precision forms of the linear solve (FP32 / BF16 block storage, FP32
BiCGSTAB + SGS inside an FP64 defect correction) against the FP64 oracle.

What is pinned, with the tolerance of each check:
  * storage rounding: FP32 storage returns exactly float32(v); BF16 storage
    returns exactly torch's float32 -> bfloat16 conversion of float32(v)
    (round to nearest, ties to even), including ties, NaN, infinities and
    overflow -- an independent implementation of the same rounding;
  * the stored blocks of an assembled Jacobian equal the rounded FP64 oracle
    Jacobian: assembly is FP64 (1e-12 of the block maximum from the oracle),
    only the store rounds, so entries agree to one unit in the last stored
    place and more than 99.9 % of them exactly;
  * defect-correction solve: the FP64 residual of the stored (rounded)
    operator, recomputed by the oracle's SpMV, meets the requested tolerance
    ||b - A_s x|| <= 1.05 rel_tol ||b|| (1.05: the oracle reduces in another
    order);
  * distance to the FP64 solution of the FP64 system: ||x - x64|| / ||x64||
    <= 1e-5 for FP32 storage (unit roundoff 6e-8 times the conditioning) and
    <= 5e-2 for BF16 storage (unit roundoff 3.9e-3);
  * FP64 form unchanged: with the default form set explicitly, the solve is
    bit-identical to a context that never called cfdg_set_precision.
"""
import numpy as np
import pytest

import cfdg
import cfdg.cases as cases
from helpers import perturbed_state, small_cases
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

CASES = small_cases()
# the SA geometries of the precision configurations: the supersonic ramp with
# no-slip walls at Re_L 5e6 and the transonic tetrahedral bump, frozen module
CASES["ramp_sa"] = cases.ramp_sa()
CASES["bump_sa"] = cases.bump_sa(ni=8, nj=8, nk=8)
FP32 = cfdg.Precision.make(cfdg.PREC_FP32, cfdg.PREC_FP32, max_outer=20, inner_rel_tol=0.1)
BF16 = cfdg.Precision.make(cfdg.PREC_BF16, cfdg.PREC_FP32, max_outer=40, inner_rel_tol=0.1)


def _system(name):
    case = CASES[name]
    mesh = cfdg.Mesh.generate(**case.mesh)
    gpu = cfdg.Solver(mesh, case.numerics, case.bcs)
    ora = Oracle(mesh, case.numerics, case.bcs)
    U = perturbed_state(ora, case)
    if name.endswith("_sa"):
        gpu.init_uniform(case.rho, case.velocity, case.p)
        turb = cases.turbulence(case, update=cfdg.TURB_FROZEN)
        gpu.set_turbulence(turb)
        ora.set_turbulence(turb)
    for s in (gpu, ora):
        s.set_state(U)
        s.eval_residual()
        s.compute_dt()
        s.assemble_jacobian()
    return mesh, gpu, ora


def _bf16_torch(v):
    torch = pytest.importorskip("torch")
    t = torch.from_numpy(np.asarray(v, dtype=np.float64).astype(np.float32)).to(torch.bfloat16)
    return t.to(torch.float32).numpy().astype(np.float64)


def test_storage_rounding_fp32_and_bf16():
    mesh, gpu, ora = _system("ramp")
    nnzb = len(gpu.bsr_pattern()[1])
    rng = np.random.default_rng(3)
    v = rng.standard_normal(nnzb * 25) * 10.0 ** rng.integers(-30, 30, nnzb * 25)
    # exact bf16 ties (1 + 2^-8 rounds down to even, 1 + 3 * 2^-8 up), signs,
    # overflow, infinities, NaN, float32 subnormals
    special = [1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8), 3.3895313892515355e38, 3.4e38,
               np.inf, -np.inf, np.nan, 1.0e-40, -3.0e-42, 0.0, -0.0]
    v[: len(special)] = special
    gpu.set_precision(FP32)
    gpu.set_bsr_values(v)
    got = gpu.bsr_values().ravel()
    exp = v.astype(np.float32).astype(np.float64)
    assert np.array_equal(got, exp, equal_nan=True)
    gpu.set_precision(BF16)
    gpu.set_bsr_values(v)
    got = gpu.bsr_values().ravel()
    exp = _bf16_torch(v)
    assert np.array_equal(np.isnan(got), np.isnan(exp))
    ok = ~np.isnan(exp)
    assert np.array_equal(got[ok], exp[ok])
    assert np.array_equal(np.signbit(got[ok]), np.signbit(exp[ok]))


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("form", ["fp32", "bf16"])
def test_stored_jacobian_is_rounded_oracle_jacobian(name, form):
    mesh, gpu, ora = _system(name)
    gpu.set_precision(FP32 if form == "fp32" else BF16)
    gpu.assemble_jacobian()
    got = gpu.bsr_values().ravel()
    ref = ora.bsr_values().ravel()
    exp = ref.astype(np.float32).astype(np.float64) if form == "fp32" else _bf16_torch(ref)
    # assembly arithmetic is FP64 and matches the oracle to 1e-12 per block;
    # rounding both to the storage type agrees except where the FP64 values
    # straddle a rounding boundary, which moves one unit in the last place
    ulp = np.abs(exp) * (2.0 ** -23 if form == "fp32" else 2.0 ** -7)
    bmax = np.repeat(np.abs(ref).reshape(-1, 25).max(axis=1), 25)
    assert np.all(np.abs(got - exp) <= ulp + 2e-12 * bmax)
    assert np.mean(got == exp) > 0.999


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("form", ["fp32", "bf16"])
def test_defect_correction_solves_stored_operator(name, form):
    mesh, gpu, ora = _system(name)
    lp = cases.default_linparams(rel_tol=1e-10, abs_tol=0.0, max_iter=2000)
    x64, r64 = gpu.linear_solve(lp)
    assert r64.status == cfdg.CONVERGED
    R, _ = ora.residual()
    b = -R.T.ravel()  # AoS right-hand side of the implicit step
    gpu.set_precision(FP32 if form == "fp32" else BF16)
    gpu.assemble_jacobian()
    x, res = gpu.linear_solve(lp)
    assert res.status == cfdg.CONVERGED, (res.status, res.iterations, res.final_residual)
    A_s = gpu.bsr_values()
    ora.set_bsr_values(A_s.ravel())
    r = b - ora.spmv(x)
    assert np.linalg.norm(r) <= 1.05e-10 * np.linalg.norm(b)
    err = np.linalg.norm(x - x64) / np.linalg.norm(x64)
    assert err <= (1e-5 if form == "fp32" else 5e-2), err


def test_default_form_is_bit_identical_to_plain():
    case = CASES["plate"]
    mesh = cfdg.Mesh.generate(**case.mesh)
    a = cfdg.Solver(mesh, case.numerics, case.bcs)
    b = cfdg.Solver(mesh, case.numerics, case.bcs)
    U = perturbed_state(a, case)
    b.set_precision(cfdg.Precision.make(cfdg.PREC_FP32, cfdg.PREC_FP32))
    b.set_precision(cfdg.Precision.make())
    lp = cases.default_linparams()
    for s in (a, b):
        s.set_state(U)
        for _ in range(2):
            s.implicit_step(lp)
    assert np.array_equal(a.get_state(), b.get_state())


def test_invalid_forms_rejected():
    case = CASES["ramp"]
    mesh = cfdg.Mesh.generate(**case.mesh)
    s = cfdg.Solver(mesh, case.numerics, case.bcs)
    for bad in (cfdg.Precision.make(cfdg.PREC_FP32, cfdg.PREC_FP64),
                cfdg.Precision.make(7, cfdg.PREC_FP32),
                cfdg.Precision.make(cfdg.PREC_FP32, cfdg.PREC_FP32, max_outer=0),
                cfdg.Precision.make(cfdg.PREC_FP32, cfdg.PREC_FP32, inner_rel_tol=1.5)):
        with pytest.raises(cfdg.CfdgError):
            s.set_precision(bad)
    p = s.get_precision()
    assert (p.block_storage, p.krylov) == (cfdg.PREC_FP64, cfdg.PREC_FP64)
