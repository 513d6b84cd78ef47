"""ctypes binding of the host front end (include/cfdg_front.h): the solver's
configuration files, its mesh container and its run driver, as the
reference's apps/solver.cpp uses them."""
from __future__ import annotations

import ctypes as C

from . import Bc, CfdgError, LinParams, Mesh, Numerics, OK, lib

_d, _i32, _i64 = C.c_double, C.c_int32, C.c_int64

# SolverConfig enumerations in the reference's order (config.hpp, linalg/config.hpp)
FLOW_INVISCID, FLOW_VISCOUS = 0, 1
SCHEME_FORWARD_EULER, SCHEME_BACKWARD_EULER = 0, 1
MODE_STEADY, MODE_UNSTEADY = 0, 1
RECON_FIRST_ORDER, RECON_MUSCL, RECON_MUSCL_DIRECTIONAL = 0, 1, 2
LIMITER_NONE, LIMITER_VENKATAKRISHNAN, LIMITER_BARTH_JESPERSEN, LIMITER_VAN_ALBADA = 0, 1, 2, 3
LIMITER_MINMOD_1D, LIMITER_VAN_ALBADA_1D = 4, 5
GRADIENT_GREEN_GAUSS_FACE, GRADIENT_GREEN_GAUSS_CELL = 0, 1
GRADIENT_LEAST_SQUARES_CELL_FACE, GRADIENT_LEAST_SQUARES_CELL_NODE = 2, 3
PRECOND_NONE, PRECOND_SGS = 0, 1
VERBOSITY_SILENT, VERBOSITY_SUMMARY, VERBOSITY_VERBOSE = 0, 1, 2


class SolverConfig(C.Structure):
    _fields_ = [
        ("eos", _i32), ("gamma", _d), ("gas_constant", _d), ("prandtl", _d), ("flow_model", _i32),
        ("init_rho", _d), ("init_p", _d), ("init_velocity", _d * 3),
        ("gradient", _i32), ("flux", _i32), ("reconstruction", _i32), ("limiter", _i32), ("venkat_k", _d),
        ("turbulence", _i32), ("nu_inf_ratio", _d), ("max_distance_sweeps", _i32), ("distance_tolerance", _d),
        ("scheme", _i32), ("cfl", _d), ("max_iterations", _i64), ("residual_tolerance", _d), ("time_mode", _i32),
        ("dt", _d), ("max_time_steps", _i64), ("bdf_order", _i64),
        ("solver", _i32), ("preconditioner", _i32), ("rel_tol", _d), ("abs_tol", _d), ("max_iter", _i32),
        ("verbosity", _i32),
        ("output_dir", C.c_char * 1024), ("field_interval", _i64), ("residual_interval", _i64),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if name == "output_dir":
                v = v.decode()
            elif name == "init_velocity":
                v = list(v)
            out[name] = v
        return out


class ConfigError(CfdgError):
    """A configuration file the reference would reject; str() carries its message."""


def _err():
    return C.create_string_buffer(4096)


def read_solver_config(path, text=None):
    """parse_solver_config of a file, or of `text` reported as `path`."""
    cfg = SolverConfig()
    e = _err()
    if text is None:
        rc = lib.cfdg_config_read_solver(str(path).encode(), C.byref(cfg), e, len(e))
    else:
        rc = lib.cfdg_config_read_solver_text(text.encode(), str(path).encode(), C.byref(cfg), e, len(e))
    if rc != OK:
        raise ConfigError(rc, e.value.decode())
    return cfg


def read_boundary_config(path, patch_names=None, mesh=None, text=None):
    """parse_boundary_config against a mesh's patches (or their names).
    Returns (list of Bc, list of warning lines)."""
    e = _err()
    w = C.create_string_buffer(1 << 16)
    if mesh is not None and text is None:
        n = mesh.n_patches
        bcs = (Bc * max(n, 1))()
        rc = lib.cfdg_config_read_boundary(str(path).encode(), mesh._h, bcs, n, w, len(w), e, len(e))
    else:
        names = list(patch_names if patch_names is not None else mesh.patch_names())
        n = len(names)
        bcs = (Bc * max(n, 1))()
        arr = (C.c_char_p * max(n, 1))(*[s.encode() for s in names])
        if text is None:
            with open(path, encoding="utf-8") as f:
                text = f.read()
        rc = lib.cfdg_config_read_boundary_text(text.encode(), str(path).encode(), arr, n, bcs, w, len(w), e, len(e))
    if rc != OK:
        raise ConfigError(rc, e.value.decode())
    warnings = [s for s in w.value.decode().split("\n") if s]
    return [bcs[i] for i in range(n)], warnings


def to_numerics(cfg):
    num, lp = Numerics(), LinParams()
    e = _err()
    rc = lib.cfdg_config_to_numerics(C.byref(cfg), C.byref(num), C.byref(lp), e, len(e))
    if rc != OK:
        raise ConfigError(rc, e.value.decode())
    return num, lp


def read_container(path):
    """import_mesh_hdf5 of a one-partition container -> Mesh."""
    h = C.c_void_p()
    e = _err()
    rc = lib.cfdg_container_read(str(path).encode(), C.byref(h), e, len(e))
    if rc != OK:
        raise CfdgError(rc, e.value.decode())
    return Mesh(h)


def write_container(mesh, path):
    e = _err()
    rc = lib.cfdg_container_write(str(path).encode(), mesh._h, e, len(e))
    if rc != OK:
        raise CfdgError(rc, e.value.decode())


def run_solver(cfg, bcs, mesh, verbose=0):
    """The reference's run on the device; returns 0 (done) or 1 (non-finite residual)."""
    arr = (Bc * len(bcs))(*bcs)
    e = _err()
    rc = lib.cfdg_run_solver(C.byref(cfg), arr, len(bcs), mesh._h, verbose, e, len(e))
    if rc < 0:
        raise CfdgError(rc, e.value.decode())
    return rc
