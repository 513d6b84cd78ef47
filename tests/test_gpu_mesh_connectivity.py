"""Cell connectivity built on the device (cfdg_mesh_connectivity with
on_device) against the host build, whose BSR pattern is the reference's and
whose lists follow their definitions (test_mesh_connectivity.py): every
offset and entry equal, on small meshes, on a 162,000-tet bump and on a
3,145,728-cell plate whose vertex-stencil candidates (about 176M keys) are
sorted in two chunks.  A context created from a mesh that holds the device
connectivity has the same BSR pattern and gives the same residual, Jacobian
blocks and wall distance, bit for bit, as one that builds its own."""
import numpy as np
import pytest

import cfdg
import cfdg.cases as cases
from test_mesh_connectivity import MESHES

pytestmark = pytest.mark.gpu

BIG = {
    "tet_bump_30": dict(kind=1, ni=30, nj=30, nk=30, x0=-1.0, x1=2.0, y0=0.0, y1=1.0, z0=0.0, z1=0.5, bump_h=0.1,
                        bump_w=0.25, bump_x=0.5, jitter=0.05, seed=7, reorder=1),
    "plate_256x128x96": cases.plate(ni=256, nj=128, nk=96).mesh,
}


def _equal(a, b):
    assert sorted(a) == sorted(b)
    for k in a:
        assert a[k].dtype == b[k].dtype and np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("name", sorted(MESHES) + sorted(BIG))
def test_device_connectivity_equals_host(name):
    spec = MESHES.get(name) or BIG[name]
    m = cfdg.Mesh.generate(**spec)
    print(f"device connectivity: {name}, {m.n_cells} cells", flush=True)
    m.build_connectivity(on_device=False, node_stencil=True)
    host = m.connectivity()
    m.build_connectivity(on_device=True, node_stencil=True)
    _equal(m.connectivity(), host)
    m.build_connectivity(on_device=True, node_stencil=False)
    dev = m.connectivity()
    assert "cell_cells_node" not in dev
    for k in dev:
        assert np.array_equal(dev[k], host[k]), k


@pytest.mark.parametrize("case_name", ["plate_sa", "bump_sa"])
def test_context_from_device_connectivity(case_name):
    case = getattr(cases, case_name)(ni=12, nj=6, nk=4)
    print(f"context from device connectivity: {case_name}", flush=True)
    out = []
    for device in (False, True):
        m = cfdg.Mesh.generate(**case.mesh)
        if device:
            m.build_connectivity(on_device=True, node_stencil=True)
        s = cfdg.Solver(m, case.numerics, case.bcs)
        s.init_uniform(case.rho, case.velocity, case.p)
        U = s.get_state()
        s.set_state(U * (1.0 + 0.02 * np.random.default_rng(3).standard_normal(U.shape)))
        s.set_turbulence(cases.turbulence(case))
        s.eval_residual()
        R, lam = s.residual()
        s.compute_dt()
        s.assemble_jacobian()
        out.append((s.bsr_pattern(), R, lam, s.bsr_values(), s.turbulence()["dist"]))
        s.close()
    (p0, R0, l0, v0, d0), (p1, R1, l1, v1, d1) = out
    for a, b in zip(p0, p1):
        assert np.array_equal(a, b)
    for a, b in ((R0, R1), (l0, l1), (v0, v1), (d0, d1)):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
