"""The device's laminar flat-plate drag against Blasius (tests/plate_drag.py)."""
import pytest

import cfdg
import plate_drag as pd

pytestmark = pytest.mark.gpu


def test_blasius_drag():
    c = pd.case()
    mesh = cfdg.Mesh.generate(**c.mesh)
    print("flat plate drag, 64x32x2 cells", flush=True)
    pd.check(pd.drag_history(cfdg.Solver(mesh, c.numerics, c.bcs), c))
