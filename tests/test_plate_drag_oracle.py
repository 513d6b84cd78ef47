"""The oracle's laminar flat-plate drag against Blasius (tests/plate_drag.py).
Runs on the CPU (about 45 s)."""
import cfdg
import plate_drag as pd
from oracle_py import Oracle


def test_blasius_drag():
    c = pd.case()
    mesh = cfdg.Mesh.generate(**c.mesh)
    pd.check(pd.drag_history(Oracle(mesh, c.numerics, c.bcs), c))
