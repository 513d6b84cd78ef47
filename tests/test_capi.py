"""The C-ABI library loads without a GPU and exports every entry point that
include/*.h (cfdg.h, cfdg_front.h) declares; host-only entry points (mesh)
work without a GPU."""
import glob
import ctypes as C
import os
import re

import numpy as np

import cfdg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(cfdg_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 30
    lib = C.CDLL(cfdg.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_mesh_generation_without_gpu():
    m = cfdg.Mesh.generate(kind=0, ni=4, nj=3, nk=2, x1=1.0, y1=1.0, z1=1.0)
    assert m.n_cells == 24
    assert m.n_faces == 3 * 24 + (3 * 2 + 4 * 2 + 4 * 3)
    vol = m.array("cell_vol")
    assert abs(vol.sum() - 1.0) < 1e-12


def test_invalid_arguments_are_rejected():
    spec = cfdg.MeshSpec(kind=0, ni=0, nj=1, nk=1, x1=1, y1=1, z1=1, y_ratio=1.0, bump_w=1.0)
    h = C.c_void_p()
    assert cfdg.lib.cfdg_mesh_generate(C.byref(spec), C.byref(h)) == cfdg.ERR_MESH
    assert cfdg.lib.cfdg_create(None, None, None, 0, C.byref(h)) == cfdg.ERR_ARG
