"""The solver application (cfdg_solver, apps/solver.cpp's command line) end to
end on the device: a mesh container and the two configuration files written
from a case, run by the binary, against the same case run through the
library's API.

Pinned: the solution files the application writes (iteration, final) are
byte-identical to those of the API run of the same numerics, boundary
conditions, initial state and loop -- so reading the container, parsing both
files, mapping them onto the device numerics and the run loop change nothing
-- and its progress lines carry the API run's residual norms in the
reference's format (solver.hpp:767-779).  Cases: the viscous plate with
BACKWARD_EULER (steady), the same with dual time stepping (BDF2), the
inviscid ramp with FORWARD_EULER, and the plate with the Spalart-Allmaras
module.  A rejected configuration ends with the reference's fatal line and
exit status 1.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import cfdg
import cfdg.cases as cases
import cfdg.front as fr

pytestmark = pytest.mark.gpu

APP = os.path.join(os.path.dirname(cfdg.LIB_PATH), "cfdg_solver")
BC_NAMES = {v: k for k, v in cfdg.BC.items() if k != "NO_SLIP_WALL_ADIABATIC"}
LIMITERS = {0: "VENKAT", 1: "BARTH", 2: "VAN_ALBADA"}
GRADIENTS = {0: "GREEN_GAUSS_FACE", 1: "GREEN_GAUSS_CELL", 2: "LEAST_SQUARES_NODE"}


def f(x):
    return repr(float(x))


def vec(v):
    return "[" + ", ".join(f(x) for x in v) + "]"


def solver_toml(case, outdir, scheme="BACKWARD_EULER", iters=12, tol=1e-12, field=5, resid=4, unsteady=None,
                turbulence=False):
    n = case.numerics
    lp = case.linparams
    flow = (f'flow_model = "VISCOUS_FLOW"\nprandtl = {f(n.prandtl)}\n' if n.flow_model == cfdg.VISCOUS_FLOW
            else 'flow_model = "INVISCID_FLOW"\n')
    numerics = 'flux = "HLLC"\n'
    if n.reconstruction == cfdg.MUSCL:
        numerics += f'reconstruction = "MUSCL"\nlimiter = "{LIMITERS[n.limiter]}"\nvenkat_k = {f(n.venkat_k)}\n'
    else:
        numerics += 'reconstruction = "FIRST_ORDER"\n'
    numerics += f'gradient = "{GRADIENTS[n.gradient]}"\n'
    time = f'scheme = "{scheme}"\ncfl = {f(n.cfl)}\nmax_iterations = {iters}\nresidual_tolerance = {f(tol)}\n'
    if unsteady:
        dt, steps, order = unsteady
        time += f'mode = "Unsteady"\ntime_step = {f(dt)}\nmax_time_steps = {steps}\nbdf_order = {order}\n'
    else:
        time += 'mode = "Steady"\n'
    text = (f'[flow]\neos = "IDEAL_GAS"\n{flow}gamma = {f(n.gamma)}\ngas_constant = {f(n.gas_constant)}\n\n'
            f'[initial]\nrho = {f(case.rho)}\npressure = {f(case.p)}\nvelocity = {vec(case.velocity)}\n\n'
            f'[numerics]\n{numerics}\n[time]\n{time}\n'
            f'[output]\ndirectory = "{outdir}"\nfield_interval = {field}\nresidual_interval = {resid}\n\n'
            f'[linalg]\nsolver = "BICGSTAB"\npreconditioner = "SGS"\nrel_tol = {f(lp.rel_tol)}\n'
            f'abs_tol = {f(lp.abs_tol)}\nmax_iter = {lp.max_iter}\n')
    if turbulence:
        text += '\n[turbulence]\nmodel = "SA"\nnu_inf_ratio = 3.0\n'
    return text


def bc_toml(case, names):
    out = []
    for p, b in enumerate(case.bcs):
        t = BC_NAMES[b.type]
        s = f'[[boundary_condition]]\npatch_id = {p}\nname = "{names[p]}"\ntype = "{t}"\n'
        if t in ("SUPERSONIC_INLET", "FARFIELD", "SUBSONIC_INLET"):
            s += f"p = {f(b.p)}\nt = {f(b.t)}\n"
            if b.inflow_mode == cfdg.INFLOW_VELOCITY:
                s += f"velocity = {vec(b.velocity)}\n"
            elif b.inflow_mode == cfdg.INFLOW_MACH_ANGLES:
                s += f"mach = {f(b.mach)}\nalpha_deg = {f(b.alpha_deg)}\nbeta_deg = {f(b.beta_deg)}\n"
            else:
                s += f"mach = {f(b.mach)}\ndirection = {vec(b.direction)}\n"
        elif t == "SUBSONIC_OUTLET":
            s += f"p_back = {f(b.p)}\n"
        elif t == "NO_SLIP_WALL":
            s += f"t_wall = {f(b.t)}\nvelocity = {vec(b.velocity)}\n"
        elif t == "NO_SLIP_WALL_HEAT_FLUX":
            s += f"tmp_grad = {f(b.tmp_grad)}\nvelocity = {vec(b.velocity)}\n"
        out.append(s)
    return "\n".join(out)


def run_app(tmp_path, case, mesh, **kw):
    fr.write_container(mesh, str(tmp_path / "mesh.h5"))
    (tmp_path / "solver.toml").write_text(solver_toml(case, "app_out", **kw), encoding="utf-8")
    (tmp_path / "bc.toml").write_text(bc_toml(case, mesh.patch_names()), encoding="utf-8")
    print(f"app: {case.name} {mesh.n_cells} cells {kw}", flush=True)
    p = subprocess.run([APP, "mesh.h5", "solver.toml", "bc.toml"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr
    return p.stderr


def same_file(a, b):
    with open(a, "rb") as fa, open(b, "rb") as fb:
        return fa.read() == fb.read()


def api_solver(case, mesh):
    s = cfdg.Solver(mesh, case.numerics, case.bcs)
    s.init_uniform(case.rho, case.velocity, case.p)
    return s


def progress(stderr):
    """(iteration, L2[mass], L2[mom], L2[energy], rel) of every progress line."""
    pat = re.compile(r"^iter (\d+)/(\d+)  wall +[\d.]+s  L2\[mass\]=(\S+)  L2\[mom\]=(\S+)  L2\[energy\]=(\S+)  "
                     r"rel=(\S+)  dt=\[(\S+),(\S+)\]$")
    out = []
    for ln in stderr.splitlines():
        m = pat.match(ln)
        if m:
            out.append((int(m.group(1)), m.group(3), m.group(4), m.group(5), m.group(6)))
    return out


def test_steady_plate(tmp_path):
    case = cases.plate(ni=24, nj=12, nk=4)
    mesh = cfdg.Mesh.generate(**case.mesh)
    err = run_app(tmp_path, case, mesh)
    s = api_solver(case, mesh)
    lp = cases.default_linparams()
    l2s = []
    for it in range(1, 13):
        l2s.append(np.array(s.implicit_step(lp).l2))
        if it % 5 == 0:
            s.write_solution_vtu(str(tmp_path / "api_out"), f"iter_{it:09d}")
    s.write_solution_vtu(str(tmp_path / "api_out"), "final_000000012")
    for stem in ("iter_000000005", "iter_000000010", "final_000000012"):
        for suffix in ("_00000.vtu", ".pvtu"):
            name = stem + suffix
            assert same_file(tmp_path / "app_out" / name, tmp_path / "api_out" / name), name
    lines = progress(err)
    assert [ln[0] for ln in lines] == [1, 4, 8, 12]
    for it, mass, mom, energy, rel in lines:
        l2 = l2s[it - 1]
        assert mass == f"{l2[0]:.4e}" and energy == f"{l2[4]:.4e}"
        assert mom == f"{np.sqrt(l2[1] ** 2 + l2[2] ** 2 + l2[3] ** 2):.4e}"
        r = 1.0 if it == 1 else max(l2[v] / max(l2s[0][v], 1e-300) for v in range(5))
        assert rel == f"{r:.3e}"
    assert "solver: done in 12 iterations" in err
    assert err.splitlines()[0].startswith("solver: mesh=mesh.h5 config=solver.toml bc=bc.toml ranks=1")


def test_dual_time_plate(tmp_path):
    case = cases.plate(ni=20, nj=10, nk=4)
    mesh = cfdg.Mesh.generate(**case.mesh)
    err = run_app(tmp_path, case, mesh, iters=8, tol=1e-2, field=2, unsteady=(1e-4, 3, 2))
    s = api_solver(case, mesh)
    s.set_unsteady(1e-4, 2)
    lp = cases.default_linparams()
    for k in range(1, 4):
        s.run_unsteady(lp, 1, 8, 1e-2)
        if k % 2 == 0:
            s.write_solution_vtu(str(tmp_path / "api_out"), f"unsteady_{k:09d}")
    s.write_solution_vtu(str(tmp_path / "api_out"), "final_000000003")
    for stem in ("unsteady_000000002", "final_000000003"):
        assert same_file(tmp_path / "app_out" / (stem + "_00000.vtu"), tmp_path / "api_out" / (stem + "_00000.vtu"))
    steps = [ln for ln in err.splitlines() if ln.startswith("Time step ")]
    assert len(steps) == 3 and steps[0].startswith("Time step 1/3: t = 1.00000e-04 s | subiters: ")


def test_forward_euler_ramp(tmp_path):
    case = cases.ramp(ni=24, nj=8, nk=2, cfl=0.5)
    mesh = cfdg.Mesh.generate(**case.mesh)
    run_app(tmp_path, case, mesh, scheme="FORWARD_EULER", iters=10, field=0, resid=5)
    s = api_solver(case, mesh)
    for _ in range(10):
        s.explicit_step()
    s.write_solution_vtu(str(tmp_path / "api_out"), "final_000000010")
    assert same_file(tmp_path / "app_out" / "final_000000010_00000.vtu",
                     tmp_path / "api_out" / "final_000000010_00000.vtu")


def test_spalart_allmaras_plate(tmp_path):
    case = cases.plate_sa(ni=20, nj=10, nk=4)
    mesh = cfdg.Mesh.generate(**case.mesh)
    run_app(tmp_path, case, mesh, iters=6, field=0, resid=3, turbulence=True)
    s = api_solver(case, mesh)
    s.set_turbulence(cases.turbulence(case))
    lp = cases.default_linparams()
    for _ in range(6):
        s.implicit_step(lp)
    s.write_solution_vtu(str(tmp_path / "api_out"), "final_000000006")
    a = tmp_path / "app_out" / "final_000000006_00000.vtu"
    assert same_file(a, tmp_path / "api_out" / "final_000000006_00000.vtu")
    assert b'Name="nu_tilde"' in a.read_bytes() and b'Name="mu_t"' in a.read_bytes()


def test_rejected_configuration(tmp_path):
    case = cases.ramp(ni=8, nj=4, nk=2)
    mesh = cfdg.Mesh.generate(**case.mesh)
    fr.write_container(mesh, str(tmp_path / "mesh.h5"))
    (tmp_path / "solver.toml").write_text(solver_toml(case, "o").replace("gamma", "gama"), encoding="utf-8")
    (tmp_path / "bc.toml").write_text(bc_toml(case, mesh.patch_names()), encoding="utf-8")
    p = subprocess.run([APP, "mesh.h5", "solver.toml", "bc.toml"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 1
    assert "[Rank 0] CGNS Reader Fatal Error: config: solver.toml [flow]: missing required key 'gamma'" in p.stderr
    p = subprocess.run([APP, "mesh.h5"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert p.returncode == 1 and "usage: cfdg_solver <mesh.h5> <solver.toml> <bc.toml> [-v|-vv]" in p.stderr
