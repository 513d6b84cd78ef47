"""Shared helpers: case construction, perturbed states, comparisons."""
import numpy as np

import cfdg
import cfdg.cases as cases


def small_cases():
    """Parity-size versions of the workloads (a few thousand cells)."""
    return {
        "ramp": cases.ramp(),
        "plate": cases.plate(ni=24, nj=16, nk=8),
        "bump": cases.bump(ni=8, nj=8, nk=8),
    }


def perturbed_state(solver_like, case, seed=1, amp=0.02):
    """Freestream with seeded multiplicative noise (deterministic)."""
    solver_like.init_uniform(case.rho, case.velocity, case.p)
    U = solver_like.get_state()
    rng = np.random.default_rng(seed)
    return U * (1.0 + amp * rng.standard_normal(U.shape))


def rel_err(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def assert_close_scaled(a, b, scale, tol, what):
    """|a - b| <= tol * scale elementwise."""
    d = np.abs(np.asarray(a) - np.asarray(b))
    bad = d > tol * np.asarray(scale)
    assert not bad.any(), f"{what}: {bad.sum()} entries exceed {tol} x scale (max ratio {np.max(d / np.maximum(scale, 1e-300)):.3e})"
