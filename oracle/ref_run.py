"""Runs the reference solver's own compiled code (oracle/_ref/ref_harness) on a
mesh + state and returns its intermediate arrays.

TEST INFRASTRUCTURE ONLY.  The binary is built by oracle/build_ref.sh from
the reference tree in the development container; where it is absent (e.g. on
a machine without the reference sources) available() is False and callers
fall back to the committed golden fixtures.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

from recio import read_records, write_records

_HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(_HERE, "_ref", "ref_harness")


def available():
    return os.path.exists(HARNESS)


def case_records(mesh_arrays, n_patches, numerics, bcs, U, init, linparams, steps=0, solver_path=1, rcm=0,
                 validate=1, geometry=0):
    """Input record dict for the harness from cfdg-style arguments."""
    rec = dict(mesh_arrays)
    rec["n_patches"] = np.array([n_patches], np.int32)
    rec["flow_model"] = np.array([numerics.flow_model], np.int32)
    rec["gamma"] = np.array([numerics.gamma])
    rec["gas_constant"] = np.array([numerics.gas_constant])
    rec["prandtl"] = np.array([numerics.prandtl])
    rec["reconstruction"] = np.array([numerics.reconstruction], np.int32)
    rec["limiter"] = np.array([numerics.limiter], np.int32)
    rec["venkat_k"] = np.array([numerics.venkat_k])
    rec["gradient"] = np.array([numerics.gradient], np.int32)
    rec["cfl"] = np.array([numerics.cfl])
    rec["bc_type"] = np.array([b.type for b in bcs], np.int32)
    rec["bc_inflow_mode"] = np.array([b.inflow_mode for b in bcs], np.int32)
    rec["bc_scalars"] = np.array([[b.p, b.t, b.tmp_grad, b.mach, b.alpha_deg, b.beta_deg] for b in bcs]).ravel()
    rec["bc_velocity"] = np.array([list(b.velocity) for b in bcs]).ravel()
    rec["bc_direction"] = np.array([list(b.direction) for b in bcs]).ravel()
    rec["U"] = np.ascontiguousarray(U, dtype=np.float64).ravel()
    rec["init"] = np.asarray(init, dtype=np.float64)
    rec["precond"] = np.array([linparams.preconditioner], np.int32)
    rec["rel_tol"] = np.array([linparams.rel_tol])
    rec["abs_tol"] = np.array([linparams.abs_tol])
    rec["max_iter"] = np.array([linparams.max_iter], np.int32)
    rec["steps"] = np.array([steps], np.int32)
    rec["solver_path"] = np.array([solver_path], np.int32)
    rec["rcm"] = np.array([rcm], np.int32)
    rec["validate"] = np.array([validate], np.int32)
    rec["geometry"] = np.array([geometry], np.int32)
    return rec


def run(records, timeout=600):
    if not available():
        raise RuntimeError("reference harness not built (oracle/build_ref.sh)")
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.rec"), os.path.join(d, "out.rec")
        write_records(fin, records)
        p = subprocess.run([HARNESS, fin, fout], capture_output=True, text=True, timeout=timeout,
                           env={**os.environ, "MPIR_CVAR_CH3_NOLOCAL": "1"})
        if p.returncode != 0:
            raise RuntimeError(f"ref_harness failed:\n{p.stdout[-2000:]}\n{p.stderr[-2000:]}")
        out = read_records(fout)
        out["_stdout"] = p.stdout
        return out


def time_residual(mesh_arrays, n_patches, numerics, bcs, U, reps=3, timeout=900):
    """Seconds per Solver::evaluate_residual of the reference on this mesh/state."""
    from cfdg import LinParams
    rec = case_records(mesh_arrays, n_patches, numerics, bcs, U, [1.0, 0.0, 0.0, 0.0, 1.0], LinParams(),
                       steps=0, solver_path=0, rcm=0, validate=0)
    rec["time_reps"] = np.array([reps], np.int32)
    out = run(rec, timeout=timeout)
    return out["time_residual"], float(out["time_setup"][0])
