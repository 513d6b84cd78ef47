"""The device's steady state of Mach 2.5 flow over a 15 degree wedge against
the oblique-shock relations (tests/shock_wedge.py)."""
import pytest

import cfdg
import shock_wedge as sw

pytestmark = pytest.mark.gpu


def test_oblique_shock_state():
    c = sw.case()
    mesh = cfdg.Mesh.generate(**c.mesh)
    print("oblique shock wedge, 48x48 cells", flush=True)
    U = sw.run(cfdg.Solver(mesh, c.numerics, c.bcs), c)
    sw.check(mesh, c, U)
