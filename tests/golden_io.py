"""Loading the golden fixtures of tests/golden/*.npz into cfdg objects."""
import glob
import os

import numpy as np

import cfdg
from cfdg import Bc, LinParams, Numerics

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, "*.npz"))
                  if not os.path.basename(p).startswith(("rcm_", "time_", "vtu_", "plate_mut", "linsys_", "geometry")))


def load(name):
    return dict(np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False))


def mesh_of(fx):
    arrays = {k[5:]: v for k, v in fx.items() if k.startswith("mesh_")}
    return cfdg.Mesh.from_arrays(arrays, int(fx["n_patches"][0]))


def numerics_of(fx):
    n = Numerics()
    for k, t in n._fields_:
        v = fx[f"num_{k}"][0]
        setattr(n, k, int(v) if t.__name__.startswith("c_int") else float(v))
    return n


def bcs_of(fx):
    out = []
    for i in range(len(fx["bc_type"])):
        b = Bc()
        b.type = int(fx["bc_type"][i])
        b.inflow_mode = int(fx["bc_inflow_mode"][i])
        b.p, b.t, b.tmp_grad, b.mach, b.alpha_deg, b.beta_deg = (float(x) for x in fx["bc_scalars"][i])
        for k in range(3):
            b.velocity[k] = float(fx["bc_velocity"][i][k])
            b.direction[k] = float(fx["bc_direction"][i][k])
        out.append(b)
    return out


def linparams_of(fx):
    lp = LinParams()
    pre, order, rel, ab, mx = fx["lin"]
    lp.preconditioner, lp.sgs_order, lp.rel_tol, lp.abs_tol, lp.max_iter = int(pre), int(order), rel, ab, int(mx)
    return lp
