"""Cell and face metrics of the mesh generator (cfdg_mesh_generate) against
the reference's own compute_mesh_geometry (src/mesh/geometry.cpp:301-539,
compiled from the reference tree into oracle/_ref/ref_harness and run by
tests/golden/make_golden.py geometry on the generator's nodes and
connectivity alone).

tests/golden/geometry.npz holds, for three meshes (jittered stretched
hexahedra with non-planar faces, renumbered tetrahedra over a bump, and
renumbered hexahedra with a split wall patch), the topology the generator
made and the reference's centroids, volumes, unit normals and areas of it.
The generator must make the same topology and the same metrics bit for bit,
so every later stage starts from the reference's geometry.  Independent of
the reference: the cells tile the box (volumes sum to its volume), and the
area vectors of each closed cell sum to zero.  Runs on the CPU.
"""
import numpy as np
import pytest

import cfdg
from golden_io import load

MESHES = {
    "hex_jitter": dict(kind=0, ni=7, nj=5, nk=4, x0=-0.3, x1=1.7, y0=0.0, y1=0.8, z0=0.0, z1=0.6, y_ratio=1.3,
                       jitter=0.2, seed=5, reorder=0),
    "tet_bump": dict(kind=1, ni=5, nj=4, nk=3, x0=0.0, x1=3.0, y0=0.0, y1=1.0, z0=0.0, z1=0.5, bump_h=0.08,
                     bump_w=0.4, bump_x=1.5, jitter=0.1, seed=2, reorder=1),
    "hex_split_rcm": dict(kind=0, ni=8, nj=4, nk=2, x0=0.0, x1=2.0, y0=0.0, y1=0.5, z0=0.0, z1=0.1, y_ratio=1.6,
                          split_ymin=1, split_ymin_x=0.5, reorder=1),
}
TOPOLOGY = ("cell_type", "cell_nodes_off", "cell_nodes", "node_x", "node_y", "node_z", "face_owner", "face_nodes_off",
            "face_nodes")
METRICS = ("cell_cx", "cell_cy", "cell_cz", "cell_vol", "face_cx", "face_cy", "face_cz", "face_nx", "face_ny", "face_nz",
           "face_area")


@pytest.fixture(scope="module")
def fx():
    return load("geometry")


@pytest.mark.parametrize("name", sorted(MESHES))
def test_metrics_match_reference(fx, name):
    m = cfdg.Mesh.generate(**MESHES[name])
    for k in TOPOLOGY:
        assert np.array_equal(m.array(k), fx[f"{name}_mesh_{k}"]), k
    for k in METRICS:
        a, b = m.array(k), fx[f"{name}_ref_{k}"]
        assert a.dtype == np.float64 and a.shape == b.shape, k
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (k, np.max(np.abs(a - b)))


@pytest.mark.parametrize("name", sorted(MESHES))
def test_closed_cells_and_tiling(fx, name):
    """Properties any correct metric has, from the reference's arrays: the
    outward area vectors of each cell cancel, owner-side normals point away
    from the owner's centroid, and (for the flat-walled boxes) the volumes
    add up to the box's."""
    s = MESHES[name]
    owner, neigh = fx[f"{name}_mesh_face_owner"], cfdg.Mesh.generate(**s).array("face_neigh")
    nc = len(fx[f"{name}_ref_cell_vol"])
    area = fx[f"{name}_ref_face_area"]
    S = np.stack([fx[f"{name}_ref_face_n{d}"] * area for d in "xyz"], axis=1)
    net = np.zeros((nc, 3))
    np.add.at(net, owner, S)
    inner = neigh >= 0
    np.add.at(net, neigh[inner], -S[inner])
    scale = np.sqrt(area.max())
    assert np.max(np.abs(net)) < 1e-13 * scale
    cc = np.stack([fx[f"{name}_ref_cell_c{d}"] for d in "xyz"], axis=1)
    fc = np.stack([fx[f"{name}_ref_face_c{d}"] for d in "xyz"], axis=1)
    assert np.all(np.einsum("ij,ij->i", S, fc - cc[owner]) > 0)
    if s.get("bump_h", 0.0) == 0.0:
        box = (s["x1"] - s["x0"]) * (s["y1"] - s["y0"]) * (s["z1"] - s["z0"])
        assert abs(fx[f"{name}_ref_cell_vol"].sum() - box) < 1e-13 * box
