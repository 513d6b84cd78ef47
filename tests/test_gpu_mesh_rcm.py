"""Reverse Cuthill-McKee order computed on the device
(cfdg_mesh_reorder_rcm_device): on the natural-order meshes of
tests/golden/rcm_*.npz the new numbering and the re-sorted faces are the
reference's (reorder.cpp:135-188, 306-576, as test_mesh_rcm.py requires of
the host order), and on larger natural-order meshes (jittered hexahedra
with a split patch, a 162,000-tet bump, the 2.1M-cell plate) the device
order equals the host order entry for entry."""
import numpy as np
import pytest

import cfdg
import cfdg.cases as cases
from golden_io import load

pytestmark = pytest.mark.gpu

BIG = {
    "hex_split_40x20x10": dict(kind=0, ni=40, nj=20, nk=10, x0=0.0, x1=2.0, y0=0.0, y1=0.5, z0=0.0, z1=0.3,
                               y_ratio=1.2, jitter=0.2, seed=3, split_ymin=1, split_ymin_x=0.5, reorder=0),
    "tet_bump_30": dict(kind=1, ni=30, nj=30, nk=30, x0=-1.0, x1=2.0, y0=0.0, y1=1.0, z0=0.0, z1=0.5, bump_h=0.1,
                        bump_w=0.25, bump_x=0.5, jitter=0.05, seed=7, reorder=0),
    "plate_256x128x64": dict(cases.plate().mesh, reorder=0),
}


@pytest.mark.parametrize("name", ["rcm_bump", "rcm_plate"])
def test_device_rcm_matches_reference(name):
    fx = load(name)
    arrays = {k[5:]: v for k, v in fx.items() if k.startswith("mesh_")}
    mesh = cfdg.Mesh.from_arrays(arrays, int(fx["mesh_face_patch"].max()) + 1)
    print(f"device rcm: {name}, {mesh.n_cells} cells", flush=True)
    mesh.reorder_rcm(on_device=True)
    assert np.array_equal(mesh.cell_gid(), fx["rcm_new2old"])
    for k in ("face_owner", "face_neigh", "face_patch", "face_area", "cell_cx"):
        assert np.array_equal(mesh.array(k), fx[f"rcm_{k}"]), k


@pytest.mark.parametrize("name", sorted(BIG))
def test_device_rcm_equals_host(name):
    host = cfdg.Mesh.generate(**BIG[name])
    dev = cfdg.Mesh.generate(**BIG[name])
    print(f"device rcm: {name}, {host.n_cells} cells", flush=True)
    host.reorder_rcm()
    dev.reorder_rcm(on_device=True)
    assert np.array_equal(dev.cell_gid(), host.cell_gid())
    for k in ("face_owner", "face_neigh", "face_patch", "cell_nodes"):
        assert np.array_equal(dev.array(k), host.array(k)), k
