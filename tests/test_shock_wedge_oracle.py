"""The oracle's steady state of Mach 2.5 flow over a 15 degree wedge against
the oblique-shock relations (tests/shock_wedge.py).  Runs on the CPU."""
import cfdg
import shock_wedge as sw
from oracle_py import Oracle


def test_oblique_shock_state():
    c = sw.case()
    mesh = cfdg.Mesh.generate(**c.mesh)
    U = sw.run(Oracle(mesh, c.numerics, c.bcs), c)
    sw.check(mesh, c, U)
