"""The cell connectivity of cfdg_mesh_connectivity on the host (the builders
the context uses): the reference's own BSR pattern (bsr_helpers.hpp:16-59,
the ref_row_ptr / ref_cols / ref_diag arrays of the golden fixtures, written
by the reference-compiled harness) and, for the lists the fixtures do not
hold, their definitions restated with plain Python sets
(aux_connectivity.cpp:11-187, 348-445): the faces of each cell in
ascending face index, its face neighbours ascending, and its vertex
neighbours ascending without duplicates and without the cell itself.
The device build is compared with this one in test_gpu_mesh_connectivity.py.
Runs on the CPU.
"""
import numpy as np
import pytest

import cfdg
from golden_io import fixture_names, load, mesh_of

MESHES = {
    "hex_jitter_rcm": dict(kind=0, ni=7, nj=5, nk=4, x0=-0.3, x1=1.7, y0=0.0, y1=0.8, z0=0.0, z1=0.6, y_ratio=1.3,
                           jitter=0.2, seed=5, reorder=1),
    "tet_bump": dict(kind=1, ni=5, nj=4, nk=3, x0=0.0, x1=3.0, y0=0.0, y1=1.0, z0=0.0, z1=0.5, bump_h=0.08,
                     bump_w=0.4, bump_x=1.5, jitter=0.1, seed=2, reorder=1),
    "hex_split": dict(kind=0, ni=8, nj=4, nk=2, x0=0.0, x1=2.0, y0=0.0, y1=0.5, z0=0.0, z1=0.1, y_ratio=1.6,
                      split_ymin=1, split_ymin_x=0.5, reorder=0),
}


def restated(m):
    """The connectivity from its definitions, as CSR (offsets, entries)."""
    nc = m.n_cells
    owner, neigh = m.array("face_owner"), m.array("face_neigh")
    coff, cn = m.array("cell_nodes_off"), m.array("cell_nodes")
    faces = [[] for _ in range(nc)]
    nbf = [[] for _ in range(nc)]
    for f in range(m.n_faces):
        faces[owner[f]].append(f)
        if neigh[f] >= 0:
            faces[neigh[f]].append(f)
    for f in range(m.n_inner_faces):
        nbf[owner[f]].append(int(neigh[f]))
        nbf[neigh[f]].append(int(owner[f]))
    node_cells = {}
    for c in range(nc):
        for v in cn[coff[c]:coff[c + 1]]:
            node_cells.setdefault(int(v), set()).add(c)
    nbn = []
    for c in range(nc):
        s = set()
        for v in cn[coff[c]:coff[c + 1]]:
            s |= node_cells[int(v)]
        s.discard(c)
        nbn.append(sorted(s))
    rows = [sorted(set(r) | {i}) for i, r in enumerate(nbf)]

    def csr(lists):
        off = np.concatenate([[0], np.cumsum([len(x) for x in lists])]).astype(np.int32)
        idx = np.array([e for x in lists for e in x], dtype=np.int32)
        return off, idx

    out = {}
    for name, lists in (("cell_faces", [sorted(x) for x in faces]), ("cell_cells_face", [sorted(x) for x in nbf]),
                        ("cell_cells_node", nbn), ("bsr", rows)):
        out[name + "_off"], out[name] = csr(lists)
    out["bsr_diag"] = np.array([out["bsr_off"][i] + rows[i].index(i) for i in range(nc)], dtype=np.int32)
    return out


@pytest.mark.parametrize("name", fixture_names())
def test_bsr_pattern_is_the_references(name):
    fx = load(name)
    m = mesh_of(fx)
    m.build_connectivity(on_device=False, node_stencil=False)
    c = m.connectivity()
    assert "cell_cells_node" not in c
    assert np.array_equal(c["bsr_off"], fx["ref_row_ptr"])
    assert np.array_equal(c["bsr"], fx["ref_cols"])
    assert np.array_equal(c["bsr_diag"], fx["ref_diag"])


@pytest.mark.parametrize("name", sorted(MESHES))
def test_lists_match_their_definitions(name):
    m = cfdg.Mesh.generate(**MESHES[name])
    m.build_connectivity(on_device=False, node_stencil=True)
    got, want = m.connectivity(), restated(m)
    assert sorted(got) == sorted(want)
    for k in want:
        assert got[k].dtype == np.int32 and np.array_equal(got[k], want[k]), k


def test_state_and_renumbering():
    m = cfdg.Mesh.generate(**MESHES["hex_split"])
    with pytest.raises(cfdg.CfdgError) as e:
        m.connectivity()
    assert e.value.code == cfdg.ERR_STATE
    m.build_connectivity(on_device=False, node_stencil=False)
    assert m.connectivity()["bsr_off"][-1] == m.n_cells + 2 * m.n_inner_faces
    m.reorder_rcm()  # renumbering drops the connectivity of the old numbering
    with pytest.raises(cfdg.CfdgError) as e:
        m.connectivity()
    assert e.value.code == cfdg.ERR_STATE
