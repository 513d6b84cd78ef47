"""The mesh container of the host front end (cfdg_container_write / _read).

The reference's container code needs parallel HDF5 and cannot be built here,
so no file it wrote exists: format parity with it is unpinned.  What these
tests pin:
* the layout the reference's writer and reader define
  (src/io/solver_mesh/hdf5_writer.cpp:195-389, hdf5_reader.cpp:140-461):
  every group, dataset and attribute with its HDF5 type and shape, checked
  with libhdf5's own h5dump on a written file -- root attributes, uint64
  partition tables of nprocs + 1 entries, int32 local indices, cell / face
  nodes_offsets without the trailing entry, [n, 3] doubles, 64-byte patch
  names and CGNS types, n_patches + 1 patch_face_offsets;
* round trips of hexahedral and tetrahedral meshes (renumbered and not),
  every array, the patch names, CGNS types and global ids identical;
* the reader's rejections: a file written for two ranks (the reference's
  message), a missing file, a file that is not HDF5.
Runs on the CPU (host code of libcfdg.so and libhdf5 only).
"""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import cfdg
import cfdg.front as fr

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
HDF5_LIB = os.environ.get("CFDG_HDF5_LIB", "/opt/conda/lib/libhdf5.so.103")


def meshes():
    yield "hex_rcm", cfdg.Mesh.generate(kind=0, ni=6, nj=4, nk=3, x0=0, x1=1, y0=0, y1=1, z0=0, z1=0.5, jitter=0.1,
                                        seed=3, reorder=1, split_ymin=1, split_ymin_x=0.4)
    yield "tet", cfdg.Mesh.generate(kind=1, ni=4, nj=3, nk=3, x0=0, x1=2, y0=0, y1=1, z0=0, z1=1, bump_h=0.05,
                                    bump_w=0.3, bump_x=1.0, seed=1)


@pytest.mark.parametrize("name", ["hex_rcm", "tet"])
def test_round_trip(tmp_path, name):
    m = dict(meshes())[name]
    m.set_patch(1, name="outlet_patch", cgns_type="BCOutflow")
    p = str(tmp_path / "m.h5")
    fr.write_container(m, p)
    r = fr.read_container(p)
    a, b = m.arrays(), r.arrays()
    for k in a:
        assert a[k].dtype == b[k].dtype and np.array_equal(a[k], b[k]), k
    assert (r.n_cells, r.n_faces, r.n_inner_faces, r.n_patches) == (m.n_cells, m.n_faces, m.n_inner_faces, m.n_patches)
    assert r.patch_names() == m.patch_names() and r.patch_names()[1] == "outlet_patch"
    assert r.patch_cgns_types() == m.patch_cgns_types() and r.patch_cgns_types()[1] == "BCOutflow"
    assert np.array_equal(r.cell_gid(), m.cell_gid())
    # a second write of what was read gives the same file contents
    p2 = str(tmp_path / "m2.h5")
    fr.write_container(r, p2)
    if os.path.exists(H5DUMP):
        d1 = subprocess.run([H5DUMP, p], capture_output=True, text=True, check=True).stdout
        d2 = subprocess.run([H5DUMP, p2], capture_output=True, text=True, check=True).stdout
        assert d1.replace(p, "F") == d2.replace(p2, "F")


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not available")
def test_layout_matches_the_reference_writer(tmp_path):
    m = dict(meshes())["hex_rcm"]
    p = str(tmp_path / "m.h5")
    fr.write_container(m, p)
    dump = subprocess.run([H5DUMP, "-H", p], capture_output=True, text=True, check=True).stdout
    # (path, HDF5 type, dims) of every object, from the header dump
    objs = {}
    stack = []
    kind = name = None
    in_type = False  # inside a compound DATATYPE block (a string type)
    for ln in dump.splitlines():
        s = ln.strip()
        if in_type:
            in_type = s != "}"
            continue
        g = re.match(r'GROUP "(.*)" \{', s)
        if g:
            stack.append(g.group(1))
            continue
        d = re.match(r'(DATASET|ATTRIBUTE) "(.*)" \{', s)
        if d:
            kind, name = d.group(1), d.group(2)
            objs[(kind, "/".join(x for x in stack if x != "/") + "/" + name)] = {}
            continue
        if kind and s.startswith("DATATYPE"):
            objs[(kind, "/".join(x for x in stack if x != "/") + "/" + name)]["type"] = s.split(None, 1)[1]
            in_type = s.endswith("{")
        elif kind and s.startswith("DATASPACE"):
            objs[(kind, "/".join(x for x in stack if x != "/") + "/" + name)]["space"] = s.split(None, 1)[1]
        elif s == "}":
            if kind:
                kind = None
            elif stack:
                stack.pop()
    nc, nf, nn, npch = m.n_cells, m.n_faces, len(m.array("node_x")), m.n_patches
    ncn, nfn = len(m.array("cell_nodes")), len(m.array("face_nodes"))
    nb = nf - m.n_inner_faces
    I32, U64, U8, F64 = "H5T_STD_I32LE", "H5T_STD_U64LE", "H5T_STD_U8LE", "H5T_IEEE_F64LE"

    def sp(*d):
        dims = ", ".join(str(x) for x in d)
        return f"SIMPLE {{ ( {dims} ) / ( {dims} ) }}"

    expect = {
        ("ATTRIBUTE", "/nprocs"): ("H5T_STD_I32LE", "SCALAR"),
        ("ATTRIBUTE", "/n_cells_global"): (U64, "SCALAR"), ("ATTRIBUTE", "/n_faces_global"): (U64, "SCALAR"),
        ("ATTRIBUTE", "/n_bfaces_global"): (U64, "SCALAR"), ("ATTRIBUTE", "/n_nodes_global"): (U64, "SCALAR"),
        ("ATTRIBUTE", "/bbox_lo"): (F64, sp(3)), ("ATTRIBUTE", "/bbox_hi"): (F64, sp(3)),
        **{("DATASET", f"partition/{k}"): (U64, sp(2)) for k in (
            "cell_offsets", "cell_nodes_offsets", "face_offsets", "face_nodes_offsets", "node_offsets",
            "comm_nb_offsets")},
        ("DATASET", "partition/part2rank"): (I32, sp(1)), ("DATASET", "partition/rank2part"): (I32, sp(1)),
        ("DATASET", "cells/type"): (U8, sp(nc)), ("DATASET", "cells/gid"): (I32, sp(nc)),
        ("DATASET", "cells/donor"): (I32, sp(nc)), ("DATASET", "cells/volume"): (F64, sp(nc)),
        ("DATASET", "cells/centroid"): (F64, sp(nc, 3)), ("DATASET", "cells/nodes_offsets"): (I32, sp(nc)),
        ("DATASET", "cells/nodes"): (I32, sp(ncn)),
        ("DATASET", "faces/owner"): (I32, sp(nf)), ("DATASET", "faces/neigh"): (I32, sp(nf)),
        ("DATASET", "faces/patch"): (I32, sp(nf)), ("DATASET", "faces/type"): (U8, sp(nf)),
        ("DATASET", "faces/area"): (F64, sp(nf)), ("DATASET", "faces/normal"): (F64, sp(nf, 3)),
        ("DATASET", "faces/centroid"): (F64, sp(nf, 3)), ("DATASET", "faces/nodes_offsets"): (I32, sp(nf)),
        ("DATASET", "faces/nodes"): (I32, sp(nfn)),
        ("DATASET", "nodes/coords"): (F64, sp(nn, 3)), ("DATASET", "nodes/gid"): (I32, sp(nn)),
        ("DATASET", "comm/nb_ranks"): (I32, sp(0)), ("DATASET", "comm/send_owned_cells"): (I32, sp(0)),
        ("DATASET", "comm/recv_ghost_cells"): (I32, sp(0)), ("DATASET", "comm/send_offsets"): (I32, sp(1)),
        ("DATASET", "comm/recv_offsets"): (I32, sp(1)),
        ("DATASET", "patches/patch_face_offsets"): (I32, sp(npch + 1)),
        ("DATASET", "patches/patch_faces"): (I32, sp(nb)),
    }
    got = {(k, n.lstrip("/") if k == "DATASET" else n): v for (k, n), v in objs.items()}
    for key, (t, s) in expect.items():
        assert key in got, key
        assert got[key]["type"] == t and got[key]["space"] == s, (key, got[key])
    for key in (("DATASET", "patches/names"), ("DATASET", "patches/cgns_types")):
        assert got[key]["space"] == sp(npch)
        assert "STRSIZE 64" in dump and "H5T_STRING" in got[key]["type"]
    assert set(got) == set(expect) | {("DATASET", "patches/names"), ("DATASET", "patches/cgns_types")}


def _set_nprocs(path, value):
    """Rewrites the root attribute nprocs with libhdf5's C API (test-only)."""
    h5 = C.CDLL(HDF5_LIB)
    h5.H5open()
    h5.H5Fopen.restype = C.c_int64
    h5.H5Fopen.argtypes = [C.c_char_p, C.c_uint, C.c_int64]
    h5.H5Aopen.restype = C.c_int64
    h5.H5Aopen.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
    h5.H5Awrite.argtypes = [C.c_int64, C.c_int64, C.c_void_p]
    h5.H5Aclose.argtypes = [C.c_int64]
    h5.H5Fclose.argtypes = [C.c_int64]
    native_int = C.c_int64.in_dll(h5, "H5T_NATIVE_INT_g").value
    f = h5.H5Fopen(path.encode(), 1, 0)  # H5F_ACC_RDWR
    a = h5.H5Aopen(f, b"nprocs", 0)
    v = C.c_int(value)
    assert h5.H5Awrite(a, native_int, C.byref(v)) >= 0
    h5.H5Aclose(a)
    h5.H5Fclose(f)


def test_rejections(tmp_path):
    m = dict(meshes())["tet"]
    p = str(tmp_path / "m.h5")
    fr.write_container(m, p)
    _set_nprocs(p, 2)
    with pytest.raises(cfdg.CfdgError) as ei:
        fr.read_container(p)
    assert ei.value.code == cfdg.ERR_IO
    assert "HDF5 mesh file was partitioned for 2 ranks, but current communicator has 1 ranks." in str(ei.value)
    missing = str(tmp_path / "none.h5")
    with pytest.raises(cfdg.CfdgError, match="Failed to open HDF5 mesh file for reading: .*none.h5"):
        fr.read_container(missing)
    bad = tmp_path / "bad.h5"
    bad.write_bytes(b"not an hdf5 file")
    with pytest.raises(cfdg.CfdgError, match="Failed to open HDF5 mesh file for reading"):
        fr.read_container(str(bad))


def test_run_solver_rejects_before_the_device(tmp_path):
    """What the device path does not run is refused before any device call."""
    m = dict(meshes())["hex_rcm"]
    text = """[flow]\neos = "IDEAL_GAS"\nflow_model = "INVISCID_FLOW"\ngamma = 1.4\ngas_constant = 287.0
[initial]\nrho = 1.0\npressure = 1e5\nvelocity = [100.0, 0.0, 0.0]
[numerics]\nflux = "HLLC"\nreconstruction = "MUSCL_DIRECTIONAL"\nlimiter = "MINMOD_1D"
[time]\nmode = "Steady"\nscheme = "BACKWARD_EULER"\ncfl = 5.0\nmax_iterations = 3\nresidual_tolerance = 1e-6
[output]\ndirectory = "out"\nfield_interval = 0\nresidual_interval = 1
"""
    cfg = fr.read_solver_config("s.toml", text=text)
    bc = "\n".join(f'[[boundary_condition]]\npatch_id = {p}\ntype = "SYMMETRY"\n' for p in range(m.n_patches))
    bcs, _ = fr.read_boundary_config("bc.toml", mesh=m, text=bc)
    with pytest.raises(cfdg.CfdgError, match="MUSCL_DIRECTIONAL reconstruction is not available") as ei:
        fr.run_solver(cfg, bcs, m)
    assert ei.value.code == cfdg.ERR_CONFIG
