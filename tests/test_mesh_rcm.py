"""Reverse Cuthill-McKee renumbering (cfdg_mesh_reorder_rcm) against the
reference's own renumbering of the same natural-order meshes
(tests/golden/rcm_*.npz, reorder_local_mesh(..., RCM) compiled from
reorder.cpp:135-188, 306-576 and run by tests/golden/make_golden.py rcm).

The new numbering (cell_gid after renumbering = the reference's permuted
cell_gid), the re-sorted face arrays and the permuted cell arrays must be
identical integer for integer and bit for bit.  This numbering fixes the
face order, the per-cell summation order and the BSR block order of every
later stage, so it is the base of the bit-exact residual and pattern
parity on reordered meshes."""
import numpy as np
import pytest

import cfdg
import cfdg.cases as cases
from golden_io import load



def _plate():
    return cases.plate(ni=10, nj=6, nk=4, reorder=0)


def _bump():
    c = cases.bump(ni=4, nj=4, nk=3)
    c.mesh["reorder"] = 0
    return c


FIXTURES = {"rcm_plate": _plate, "rcm_bump": _bump}


@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_generator_reproduces_fixture_mesh(name):
    """The committed natural-order mesh is what the generator makes (so the
    renumbering below starts from the reference's input)."""
    fx = load(name)
    mesh = cfdg.Mesh.generate(**FIXTURES[name]().mesh)
    for k, a in mesh.arrays().items():
        assert np.array_equal(a, fx[f"mesh_{k}"]), k


@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_rcm_matches_reference(name):
    fx = load(name)
    arrays = {k[5:]: v for k, v in fx.items() if k.startswith("mesh_")}
    n_patches = int(fx["mesh_face_patch"].max()) + 1
    mesh = cfdg.Mesh.from_arrays(arrays, n_patches)
    assert np.array_equal(mesh.cell_gid(), np.arange(mesh.n_cells))
    mesh.reorder_rcm()
    assert np.array_equal(mesh.cell_gid(), fx["rcm_new2old"])
    assert np.array_equal(mesh.array("face_owner"), fx["rcm_face_owner"])
    assert np.array_equal(mesh.array("face_neigh"), fx["rcm_face_neigh"])
    assert np.array_equal(mesh.array("face_patch"), fx["rcm_face_patch"])
    assert np.array_equal(mesh.array("face_area"), fx["rcm_face_area"])
    assert np.array_equal(mesh.array("cell_cx"), fx["rcm_cell_cx"])
    # the permuted cell arrays follow the permutation
    n2o = fx["rcm_new2old"]
    for k in ("cell_cx", "cell_cy", "cell_cz", "cell_vol"):
        assert np.array_equal(mesh.array(k), arrays[k][n2o]), k


@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_rcm_is_a_bandwidth_reducing_permutation(name):
    """Properties any RCM order has, independent of the reference: a
    permutation, and a face-graph bandwidth no larger than the natural
    order's (here strictly smaller for the tetrahedra)."""
    fx = load(name)
    arrays = {k[5:]: v for k, v in fx.items() if k.startswith("mesh_")}
    n2o = fx["rcm_new2old"]
    assert np.array_equal(np.sort(n2o), np.arange(len(n2o)))

    def bandwidth(owner, neigh):
        inner = neigh >= 0
        return int(np.max(np.abs(owner[inner] - neigh[inner])))

    b0 = bandwidth(arrays["face_owner"], arrays["face_neigh"])
    b1 = bandwidth(fx["rcm_face_owner"], fx["rcm_face_neigh"])
    assert b1 <= b0
