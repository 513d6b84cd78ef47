"""Mesh metrics computed on the device (cfdg_mesh_geometry_device) against
the host's and the reference's: on the three meshes of
tests/golden/geometry.npz (the reference's compute_mesh_geometry,
geometry.cpp:301-539) and on larger hexahedral and tetrahedral meshes, every
centroid, volume, unit normal and area equal bit for bit; a degenerate cell
is rejected and leaves the mesh unchanged."""
import numpy as np
import pytest

import cfdg
from golden_io import load
from test_mesh_geometry import MESHES, METRICS

pytestmark = pytest.mark.gpu

BIG = {
    "hex_60x40x30": dict(kind=0, ni=60, nj=40, nk=30, x0=0.0, x1=2.0, y0=0.0, y1=1.0, z0=0.0, z1=0.7, y_ratio=1.05,
                         jitter=0.2, seed=9, reorder=1),
    "tet_bump_30": dict(kind=1, ni=30, nj=30, nk=30, x0=-1.0, x1=2.0, y0=0.0, y1=1.0, z0=0.0, z1=0.5, bump_h=0.1,
                        bump_w=0.25, bump_x=0.5, jitter=0.05, seed=7, reorder=1),
}


def _same_bits(a, b):
    return a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("name", sorted(MESHES) + sorted(BIG))
def test_device_metrics_equal_host_metrics(name):
    spec = MESHES.get(name) or BIG[name]
    m = cfdg.Mesh.generate(**spec)
    host = {k: m.array(k).copy() for k in METRICS}
    print(f"device geometry: {name}, {m.n_cells} cells", flush=True)
    m.compute_geometry_on_device()
    for k in METRICS:
        assert _same_bits(m.array(k), host[k]), k
    if name in MESHES:
        fx = load("geometry")
        for k in METRICS:
            assert _same_bits(m.array(k), fx[f"{name}_ref_{k}"]), k


def test_degenerate_cell_is_rejected():
    m = cfdg.Mesh.generate(**MESHES["hex_split_rcm"])
    arrays = m.arrays()
    before = {k: m.array(k).copy() for k in METRICS}
    # collapse the first cell: all its nodes onto its first node
    cn = arrays["cell_nodes"][arrays["cell_nodes_off"][0]:arrays["cell_nodes_off"][1]]
    for k in ("node_x", "node_y", "node_z"):
        arrays[k][cn] = arrays[k][cn[0]]
    bad = cfdg.Mesh.from_arrays(arrays, m.n_patches)
    with pytest.raises(cfdg.CfdgError) as ei:
        bad.compute_geometry_on_device()
    assert ei.value.code == cfdg.ERR_MESH
    for k in METRICS:  # left unchanged: the metrics the arrays came with
        assert np.array_equal(bad.array(k), before[k]), k
