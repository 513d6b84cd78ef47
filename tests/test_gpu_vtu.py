"""Solution files of the device path against the reference's own
(tests/golden/vtu_*.npz: the VTU piece and PVTU master Solver::run wrote after
a few implicit steps, vtu.cpp:321-475, with the state they hold).

Given the reference's state, cfdg_write_solution_vtu forms rho, u, v, w,
pressure and Mach on the device with the reference's operations and prints
them in its formats: the files must be byte-identical (hexahedra and
tetrahedra)."""
import numpy as np
import pytest

import cfdg
from golden_io import bcs_of, linparams_of, load, mesh_of, numerics_of

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["vtu_ramp", "vtu_bump"])
def test_solution_files_byte_identical(name, tmp_path):
    fx = load(name)
    mesh = mesh_of(fx)
    s = cfdg.Solver(mesh, numerics_of(fx), bcs_of(fx))
    s.set_state(fx["ref_steps_U"].reshape(5, -1))
    stem = f"final_{int(fx['steps'][0]):09d}"
    s.write_solution_vtu(tmp_path / "out", stem)
    piece = (tmp_path / "out" / f"{stem}_00000.vtu").read_bytes()
    master = (tmp_path / "out" / f"{stem}.pvtu").read_bytes()
    assert piece == bytes(fx["ref_vtu_piece"])
    assert master == bytes(fx["ref_vtu_master"])


def test_solution_file_fields_after_device_steps(tmp_path):
    """After the same implicit steps run on the device, the printed fields
    agree with the reference's to the BiCGSTAB reduction tolerance."""
    fx = load("vtu_ramp")
    mesh = mesh_of(fx)
    s = cfdg.Solver(mesh, numerics_of(fx), bcs_of(fx))
    rho, u, v, w, p = fx["init"]
    s.init_uniform(rho, (u, v, w), p)
    lp = linparams_of(fx)
    for _ in range(int(fx["steps"][0])):
        s.implicit_step(lp)
    s.write_solution_vtu(tmp_path, "dev")

    def fields(text):
        out = {}
        for name in ("rho", "u", "v", "w", "pressure", "mach"):
            key = f'Name="{name}" format="ascii">'
            body = text.split(key, 1)[1].split("</DataArray>", 1)[0]
            out[name] = np.array(body.split(), dtype=np.float64)
        return out

    got = fields((tmp_path / "dev_00000.vtu").read_text())
    ref = fields(bytes(fx["ref_vtu_piece"]).decode())
    for k in got:
        scale = max(np.abs(ref[k]).max(), 1e-300)
        assert np.abs(got[k] - ref[k]).max() <= 1e-8 * scale, k


def test_mesh_mismatch_rejected(tmp_path):
    fx, other = load("vtu_ramp"), load("vtu_bump")
    s = cfdg.Solver(mesh_of(fx), numerics_of(fx), bcs_of(fx))
    s.mesh = mesh_of(other)
    with pytest.raises(RuntimeError):
        s.write_solution_vtu(tmp_path, "x")


def test_turbulence_fields_appended(tmp_path):
    """With the turbulence module the piece and the master list rho, u, v, w,
    pressure, mach, nu_tilde, mu_t in that order (the module appends its two
    fields, spalart_allmaras.hpp:445-448; the reference's SA module does not
    compile, so the order is the one that header states); the module fields
    are those of the last residual evaluation, printed with the format of the
    mean fields."""
    import cfdg.cases as cases
    case = cases.plate_sa(ni=12, nj=8, nk=4)
    mesh = cfdg.Mesh.generate(**case.mesh)
    s = cfdg.Solver(mesh, case.numerics, case.bcs)
    s.init_uniform(case.rho, case.velocity, case.p)
    s.set_turbulence(cases.turbulence(case))
    s.write_solution_vtu(tmp_path, "before")
    before = (tmp_path / "before_00000.vtu").read_text()
    s.implicit_step(cases.default_linparams())
    s.eval_residual()
    t = s.turbulence()
    s.write_solution_vtu(tmp_path, "sa")
    text = (tmp_path / "sa_00000.vtu").read_text()
    master = (tmp_path / "sa.pvtu").read_text()
    names = ["rho", "u", "v", "w", "pressure", "mach", "nu_tilde", "mu_t"]
    pos = [text.index(f'Name="{n}" format="ascii">') for n in names]
    assert pos == sorted(pos)
    mpos = [master.index(f'<PDataArray type="Float64" Name="{n}"/>') for n in names]
    assert mpos == sorted(mpos)

    def field(txt, name):
        body = txt.split(f'Name="{name}" format="ascii">', 1)[1].split("</DataArray>", 1)[0]
        return body.split()

    assert field(text, "nu_tilde") == [f"{x:.10e}" for x in t["nu"]]
    assert field(text, "mu_t") == [f"{x:.10e}" for x in t["mut"]]
    assert set(field(before, "mu_t")) == {f"{0.0:.10e}"}
    # without the module the files keep the six mean fields only
    s2 = cfdg.Solver(mesh, case.numerics, case.bcs)
    s2.init_uniform(case.rho, case.velocity, case.p)
    s2.write_solution_vtu(tmp_path, "lam")
    assert "nu_tilde" not in (tmp_path / "lam_00000.vtu").read_text()
    assert "mu_t" not in (tmp_path / "lam.pvtu").read_text()
