"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, to compare against.  The API mirrors cfdg.Solver so
a test runs the same calls on both and compares.  Argument blocks are the
cfdg ctypes structures (identical layouts, see oracle.h).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    vp, d, i32, i64 = C.c_void_p, C.c_double, C.c_int32, C.c_int64
    sig = {
        "oracle_create": [vp, vp, vp, i32, vp], "oracle_destroy": [vp], "oracle_sizes": [vp, vp, vp, vp],
        "oracle_set_state": [vp, vp, i64], "oracle_get_state": [vp, vp, i64], "oracle_init_uniform": [vp, d, vp, d],
        "oracle_eval_residual": [vp], "oracle_get_primitives": [vp, vp, i64],
        "oracle_get_gradients": [vp, vp, i64, vp, i64], "oracle_get_residual": [vp, vp, i64, vp],
        "oracle_residual_norms": [vp, vp], "oracle_get_flux_scale": [vp, vp, i64], "oracle_compute_dt": [vp], "oracle_get_dt": [vp, vp, vp],
        "oracle_assemble_jacobian": [vp], "oracle_get_bsr_pattern": [vp, vp, vp, vp],
        "oracle_get_bsr_values": [vp, vp], "oracle_set_bsr_values": [vp, vp],
        "oracle_linear_solve": [vp, vp, vp, vp, vp, vp], "oracle_implicit_step": [vp, vp, vp],
        "oracle_row_groups": [vp, vp, vp, vp, vp, vp, vp], "oracle_sgs_apply": [vp, i32, vp, vp],
        "oracle_spmv": [vp, vp, vp], "oracle_boundary_integrals": [vp, vp, vp, vp, i32],
        "oracle_wall_distance": [vp, i32, i32, d, vp],
        "oracle_set_turbulence": [vp, vp], "oracle_set_turbulence_state": [vp, vp],
        "oracle_set_eddy_viscosity": [vp, vp, i32],
        "oracle_get_turbulence": [vp, vp, vp, vp, vp, vp, vp, vp, vp],
        "oracle_sa_point": [d, d, d, d, d, d, vp],
        "oracle_explicit_step": [vp, vp], "oracle_set_unsteady": [vp, d, i32],
        "oracle_run_unsteady": [vp, vp, i64, i64, d, vp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.oracle_last_error.argtypes = [vp]
    lib.oracle_last_error.restype = C.c_char_p
    lib.oracle_sa_constant.argtypes = [i32]
    lib.oracle_sa_constant.restype = d
    return lib


lib = _load()


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class OracleError(RuntimeError):
    pass


class Oracle:
    def __init__(self, mesh, numerics, bcs):
        """mesh: cfdg.Mesh (its MeshView is passed through unchanged)."""
        self.mesh = mesh
        bc_type = type(bcs[0])
        arr = (bc_type * len(bcs))(*bcs)
        h = C.c_void_p()
        rc = lib.oracle_create(C.byref(mesh.view), C.byref(numerics), arr, len(bcs), C.byref(h))
        if rc != 0:
            raise OracleError("oracle_create failed")
        self._h = h
        nc, nt, nz = C.c_int32(), C.c_int32(), C.c_int64()
        lib.oracle_sizes(h, C.byref(nc), C.byref(nt), C.byref(nz))
        self.n_cells, self.n_total, self.nnzb = nc.value, nt.value, nz.value
        self.n_patches = len(bcs)

    def close(self):
        if getattr(self, "_h", None):
            lib.oracle_destroy(self._h)
            self._h = None

    __del__ = close

    def _chk(self, rc):
        if rc != 0:
            raise OracleError(lib.oracle_last_error(self._h).decode())

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        self._chk(lib.oracle_set_state(self._h, _ptr(U), self.n_cells))

    def get_state(self):
        U = np.empty((5, self.n_cells))
        self._chk(lib.oracle_get_state(self._h, _ptr(U), self.n_cells))
        return U

    def init_uniform(self, rho, velocity, p):
        vel = (C.c_double * 3)(*velocity)
        self._chk(lib.oracle_init_uniform(self._h, rho, vel, p))

    def eval_residual(self):
        self._chk(lib.oracle_eval_residual(self._h))

    def primitives(self):
        q = np.empty((5, self.n_total))
        self._chk(lib.oracle_get_primitives(self._h, _ptr(q), self.n_total))
        return q

    def gradients(self):
        g = np.empty((15, self.n_total))
        phi = np.empty((5, self.n_cells))
        self._chk(lib.oracle_get_gradients(self._h, _ptr(g), self.n_total, _ptr(phi), self.n_cells))
        return g, phi

    def residual(self):
        R = np.empty((5, self.n_cells))
        lam = np.empty(self.n_cells)
        self._chk(lib.oracle_get_residual(self._h, _ptr(R), self.n_cells, _ptr(lam)))
        return R, lam

    def flux_scale(self):
        S = np.empty((5, self.n_cells))
        self._chk(lib.oracle_get_flux_scale(self._h, _ptr(S), self.n_cells))
        return S

    def residual_norms(self):
        l2 = np.empty(5)
        self._chk(lib.oracle_residual_norms(self._h, _ptr(l2)))
        return l2

    def compute_dt(self):
        self._chk(lib.oracle_compute_dt(self._h))

    def dt(self):
        dt = np.empty(self.n_cells)
        al = np.empty(self.n_cells)
        self._chk(lib.oracle_get_dt(self._h, _ptr(dt), _ptr(al)))
        return dt, al

    def assemble_jacobian(self):
        self._chk(lib.oracle_assemble_jacobian(self._h))

    def bsr_pattern(self):
        rp = np.empty(self.n_cells + 1, np.int32)
        cols = np.empty(self.nnzb, np.int32)
        dg = np.empty(self.n_cells, np.int32)
        self._chk(lib.oracle_get_bsr_pattern(self._h, _ptr(rp), _ptr(cols), _ptr(dg)))
        return rp, cols, dg

    def bsr_values(self):
        v = np.empty((self.nnzb, 5, 5))
        self._chk(lib.oracle_get_bsr_values(self._h, _ptr(v)))
        return v

    def set_bsr_values(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        self._chk(lib.oracle_set_bsr_values(self._h, _ptr(v)))

    def linear_solve(self, lp, b=None, x0=None):
        # result struct layout identical to cfdg.IterResult
        from cfdg import IterResult
        x = np.empty(5 * self.n_cells)
        res = IterResult()
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        xx = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        self._chk(lib.oracle_linear_solve(self._h, C.byref(lp), _ptr(bb), _ptr(xx), _ptr(x), C.byref(res)))
        return x, res

    def implicit_step(self, lp):
        from cfdg import StepResult
        out = StepResult()
        self._chk(lib.oracle_implicit_step(self._h, C.byref(lp), C.byref(out)))
        return out

    def explicit_step(self):
        """One forward Euler step; returns the L2 norms of the residual before the update."""
        l2 = np.zeros(5)
        self._chk(lib.oracle_explicit_step(self._h, _ptr(l2)))
        return l2

    def set_unsteady(self, dt, order):
        self._chk(lib.oracle_set_unsteady(self._h, float(dt), int(order)))

    def run_unsteady(self, lp, steps, max_subiter, sub_tol):
        """Dual time stepping; returns the sub-iterations of every physical step."""
        sub = np.zeros(int(steps), np.int64)
        self._chk(lib.oracle_run_unsteady(self._h, C.byref(lp), int(steps), int(max_subiter), float(sub_tol), _ptr(sub)))
        return sub

    def row_groups(self):
        n = self.n_cells
        col, fl, bl = np.empty(n, np.int32), np.empty(n, np.int32), np.empty(n, np.int32)
        nc, nf, nb = C.c_int32(), C.c_int32(), C.c_int32()
        self._chk(lib.oracle_row_groups(self._h, _ptr(col), C.byref(nc), _ptr(fl), C.byref(nf), _ptr(bl), C.byref(nb)))
        return col, nc.value, fl, nf.value, bl, nb.value

    def sgs_apply(self, order, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty_like(r)
        self._chk(lib.oracle_sgs_apply(self._h, order, _ptr(r), _ptr(z)))
        return z

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self._chk(lib.oracle_spmv(self._h, _ptr(x), _ptr(y)))
        return y

    def wall_distance(self, stencil=1, max_sweeps=500, rel_tol=1e-8):
        d = np.empty(self.n_cells)
        self._chk(lib.oracle_wall_distance(self._h, stencil, max_sweeps, rel_tol, _ptr(d)))
        return d

    def boundary_integrals(self):
        """Per patch: net mass flux, energy flux, force (momentum flux, 3)."""
        n = self.n_patches
        mass, energy, force = np.empty(n), np.empty(n), np.empty((n, 3))
        self._chk(lib.oracle_boundary_integrals(self._h, _ptr(mass), _ptr(energy), _ptr(force), n))
        return mass, energy, force

    def set_turbulence(self, turb):
        """turb: cfdg.Turbulence (identical layout)."""
        self._chk(lib.oracle_set_turbulence(self._h, C.byref(turb)))

    def set_turbulence_state(self, rho_nu):
        rho_nu = np.ascontiguousarray(rho_nu, dtype=np.float64)
        self._chk(lib.oracle_set_turbulence_state(self._h, _ptr(rho_nu)))

    def set_eddy_viscosity(self, mut):
        """Prescribed mu_t of every cell and ghost (n_total), or None to clear."""
        if mut is None:
            self._chk(lib.oracle_set_eddy_viscosity(self._h, None, 0))
            return
        mut = np.ascontiguousarray(mut, dtype=np.float64)
        self._chk(lib.oracle_set_eddy_viscosity(self._h, _ptr(mut), mut.size))

    def turbulence(self):
        """Fields of the last residual evaluation: dict of rho_nu, nu, mut,
        dist, res, diag (n_cells), grad (3, n_cells), mdot (n_faces)."""
        n = self.n_cells
        out = {k: np.empty(n) for k in ("rho_nu", "nu", "mut", "dist", "res", "diag")}
        out["grad"] = np.empty((3, n))
        out["mdot"] = np.empty(self.mesh.n_faces)
        self._chk(lib.oracle_get_turbulence(self._h, *(_ptr(out[k]) for k in
                                                       ("rho_nu", "nu", "mut", "dist", "res", "grad", "mdot", "diag"))))
        return out


SA_CONSTANTS = ("cb1", "cb2", "sigma", "kappa", "cw1", "cw2", "cw3", "cv1")


def sa_constants():
    return {k: lib.oracle_sa_constant(i) for i, k in enumerate(SA_CONSTANTS)}


def sa_point(rho, nu, mu_lam, d, S, grad2):
    """One cell's SA source evaluation: dict of P, D, cb2, fv1, fv2, sbar, r, fw."""
    out = np.empty(8)
    if lib.oracle_sa_point(rho, nu, mu_lam, d, S, grad2, _ptr(out)) != 0:
        raise OracleError("oracle_sa_point failed")
    return dict(zip(("P", "D", "cb2", "fv1", "fv2", "sbar", "r", "fw"), out))
