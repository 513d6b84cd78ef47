"""The reference-side binding of INTEGRATION.md compiles and binds.

The C++ code blocks of INTEGRATION.md sections 1 and 2 are extracted as
written and compiled against the reference's own headers
(include/cfd/mesh/localmesh.hpp, bc/config.hpp, solver/compressible/
config.hpp, with its enumerations held to cfdg.h / cfdg_front.h by the
blocks' static_asserts), linked with the reference's compiled configuration
parsers (oracle/_ref/obj, built by oracle/build_ref.sh) and libcfdg.so, and
run:
* a MeshPart filled from a generated mesh goes through to_cfdg and comes
  back through cfdg_mesh_get_view array for array, patch names included;
* a solver and a boundary file parsed by the reference's
  parse_solver_config / parse_boundary_config and mapped by to_cfdg_config
  give the same cfdg_numerics, cfdg_linparams and cfdg_bc values as the
  front end's cfdg_config_read_solver / _boundary and
  cfdg_config_to_numerics of the same files.
Runs on the CPU where the reference tree and its build are present.
"""
import os
import re
import shutil
import subprocess

import pytest

import cfdg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
OBJ = os.path.join(ROOT, "oracle", "_ref", "obj")
MPICXX = "/opt/conda/bin/mpicxx"
REF_OBJS = ["src_mpi_log.o", "src_io_toml_toml_utilities.o", "src_solver_compressible_config.o", "src_bc_config.o"]

pytestmark = pytest.mark.skipif(
    not (os.path.isdir(os.path.join(REF, "include")) and os.path.exists(MPICXX)
         and all(os.path.exists(os.path.join(OBJ, o)) for o in REF_OBJS)),
    reason="needs the reference tree and oracle/build_ref.sh's objects")

MAIN = r'''
#include <mpi.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#define CHECK(c) do { if (!(c)) { std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); return 1; } } while (0)

template <class A, class B>
bool same(const A* a, const std::vector<B>& b, size_t n) {
    if (b.size() < n) return false;
    for (size_t i = 0; i < n; ++i) if (static_cast<double>(a[i]) != static_cast<double>(b[i])) return false;
    return true;
}

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    const std::string dir = argv[1];
    // --- section 1: a MeshPart from a generated mesh, through to_cfdg and back
    cfdg_mesh_spec spec{};
    spec.kind = 1; spec.ni = 4; spec.nj = 3; spec.nk = 2; spec.x1 = 2.0; spec.y1 = 1.0; spec.z1 = 0.5;
    spec.y_ratio = 1.0; spec.bump_h = 0.05; spec.bump_w = 0.3; spec.bump_x = 1.0; spec.seed = 5; spec.reorder = 1;
    cfdg_mesh* g = nullptr;
    CHECK(cfdg_mesh_generate(&spec, &g) == CFDG_OK);
    cfdg_mesh_view v{};
    cfdg_mesh_get_view(g, &v);
    cfd::mesh::MeshPart mp;
    mp.nprocs = 1;
    mp.n_cells = mp.n_own = v.n_cells; mp.n_nodes = mp.n_nodes_own = v.n_nodes; mp.n_faces = v.n_faces;
    mp.n_inner_faces = mp.n_internal_faces = v.n_inner_faces;
    for (int i = 0; i < v.n_cells; ++i) {
        mp.cell_type.push_back(static_cast<cfd::mesh::CellType>(v.cell_type[i]));
        mp.cell_centroid_x.push_back(v.cell_cx[i]); mp.cell_centroid_y.push_back(v.cell_cy[i]);
        mp.cell_centroid_z.push_back(v.cell_cz[i]); mp.cell_volume.push_back(v.cell_vol[i]);
    }
    mp.cell_nodes_offsets.assign(v.cell_nodes_off, v.cell_nodes_off + v.n_cells + 1);
    mp.cell_nodes.assign(v.cell_nodes, v.cell_nodes + v.cell_nodes_off[v.n_cells]);
    mp.node_x.assign(v.node_x, v.node_x + v.n_nodes); mp.node_y.assign(v.node_y, v.node_y + v.n_nodes);
    mp.node_z.assign(v.node_z, v.node_z + v.n_nodes);
    for (int f = 0; f < v.n_faces; ++f) {
        mp.face_type.push_back(static_cast<cfd::mesh::CellType>(v.face_type[f]));
        mp.face_owner.push_back(v.face_owner[f]); mp.face_neigh.push_back(v.face_neigh[f]);
        mp.face_patch.push_back(v.face_patch[f]); mp.face_area.push_back(v.face_area[f]);
        mp.face_centroid_x.push_back(v.face_cx[f]); mp.face_centroid_y.push_back(v.face_cy[f]);
        mp.face_centroid_z.push_back(v.face_cz[f]); mp.face_normal_x.push_back(v.face_nx[f]);
        mp.face_normal_y.push_back(v.face_ny[f]); mp.face_normal_z.push_back(v.face_nz[f]);
    }
    mp.face_nodes_offsets.assign(v.face_nodes_off, v.face_nodes_off + v.n_faces + 1);
    mp.face_nodes.assign(v.face_nodes, v.face_nodes + v.face_nodes_off[v.n_faces]);
    for (int p = 0; p < v.n_patches; ++p) mp.patches.push_back({"patch_" + std::to_string(p), "BCType" + std::to_string(p)});
    cfdg_mesh* m = to_cfdg(mp);
    cfdg_mesh_view w{};
    cfdg_mesh_get_view(m, &w);
    CHECK(w.n_cells == mp.n_cells && w.n_faces == mp.n_faces && w.n_nodes == mp.n_nodes);
    CHECK(w.n_inner_faces == mp.n_inner_faces && w.n_patches == static_cast<int32_t>(mp.patches.size()));
    const size_t nc = mp.n_cells, nf = mp.n_faces, nn = mp.n_nodes;
    CHECK(same(w.cell_cx, mp.cell_centroid_x, nc) && same(w.cell_cy, mp.cell_centroid_y, nc));
    CHECK(same(w.cell_cz, mp.cell_centroid_z, nc) && same(w.cell_vol, mp.cell_volume, nc));
    CHECK(same(w.cell_nodes_off, mp.cell_nodes_offsets, nc + 1) && same(w.cell_nodes, mp.cell_nodes, mp.cell_nodes.size()));
    CHECK(same(w.node_x, mp.node_x, nn) && same(w.node_y, mp.node_y, nn) && same(w.node_z, mp.node_z, nn));
    CHECK(same(w.face_owner, mp.face_owner, nf) && same(w.face_neigh, mp.face_neigh, nf));
    CHECK(same(w.face_patch, mp.face_patch, nf) && same(w.face_area, mp.face_area, nf));
    CHECK(same(w.face_nx, mp.face_normal_x, nf) && same(w.face_ny, mp.face_normal_y, nf) && same(w.face_nz, mp.face_normal_z, nf));
    CHECK(same(w.face_cx, mp.face_centroid_x, nf) && same(w.face_cy, mp.face_centroid_y, nf) && same(w.face_cz, mp.face_centroid_z, nf));
    CHECK(same(w.face_nodes_off, mp.face_nodes_offsets, nf + 1) && same(w.face_nodes, mp.face_nodes, mp.face_nodes.size()));
    for (size_t i = 0; i < nc; ++i) CHECK(w.cell_type[i] == static_cast<uint8_t>(mp.cell_type[i]));
    for (size_t f = 0; f < nf; ++f) CHECK(w.face_type[f] == static_cast<uint8_t>(mp.face_type[f]));
    char buf[128];
    for (int p = 0; p < w.n_patches; ++p) {
        cfdg_mesh_patch_name(m, p, buf, sizeof buf);
        CHECK(mp.patches[static_cast<size_t>(p)].name == buf);
        cfdg_mesh_patch_cgns_type(m, p, buf, sizeof buf);
        CHECK(mp.patches[static_cast<size_t>(p)].cgns_type == buf);
    }
    // --- section 2: the reference's parsed files through to_cfdg_config vs the front end
    const std::string sf = dir + "/solver.toml", bf = dir + "/bc.toml";
    const auto cfg = cc::parse_solver_config(sf, MPI_COMM_WORLD);
    const auto bcfg = cfd::bc::parse_boundary_config(bf, mp, MPI_COMM_WORLD);
    cfdg_numerics num{}; cfdg_linparams lp{}; std::vector<cfdg_bc> bcs;
    CHECK(to_cfdg_config(cfg, bcfg, num, lp, bcs));
    cfdg_solver_config c{};
    char err[1024];
    CHECK(cfdg_config_read_solver(sf.c_str(), &c, err, sizeof err) == CFDG_OK);
    cfdg_numerics num2{}; cfdg_linparams lp2{};
    CHECK(cfdg_config_to_numerics(&c, &num2, &lp2, err, sizeof err) == CFDG_OK);
    CHECK(num.flow_model == num2.flow_model && num.gamma == num2.gamma && num.gas_constant == num2.gas_constant);
    CHECK(num.prandtl == num2.prandtl && num.prandtl_turb == num2.prandtl_turb);
    CHECK(num.reconstruction == num2.reconstruction && num.limiter == num2.limiter && num.venkat_k == num2.venkat_k);
    CHECK(num.gradient == num2.gradient && num.cfl == num2.cfl);
    CHECK(lp.preconditioner == lp2.preconditioner && lp.sgs_order == lp2.sgs_order && lp.rel_tol == lp2.rel_tol);
    CHECK(lp.abs_tol == lp2.abs_tol && lp.max_iter == lp2.max_iter);
    std::vector<cfdg_bc> bcs2(bcs.size());
    CHECK(cfdg_config_read_boundary(bf.c_str(), m, bcs2.data(), static_cast<int32_t>(bcs2.size()), nullptr, 0, err,
                                    sizeof err) == CFDG_OK);
    for (size_t p = 0; p < bcs.size(); ++p) {
        const cfdg_bc& a = bcs[p]; const cfdg_bc& b = bcs2[p];
        CHECK(a.type == b.type && a.inflow_mode == b.inflow_mode && a.p == b.p && a.t == b.t && a.tmp_grad == b.tmp_grad);
        CHECK(a.mach == b.mach && a.alpha_deg == b.alpha_deg && a.beta_deg == b.beta_deg);
        for (int k = 0; k < 3; ++k) CHECK(a.velocity[k] == b.velocity[k] && a.direction[k] == b.direction[k]);
    }
    cfdg_mesh_free(m);
    cfdg_mesh_free(g);
    std::printf("binding ok: %zu cells, %zu patches\n", nc, bcs.size());
    MPI_Finalize();
    return 0;
}
'''

SOLVER = """[flow]
eos = "IDEAL_GAS"
flow_model = "VISCOUS_FLOW"
prandtl = 0.72
gamma = 1.4
gas_constant = 287.05
[initial]
rho = 1.2
pressure = 101325
velocity = [50.0, 1.0, 0.0]
[numerics]
flux = "HLLC"
reconstruction = "MUSCL"
limiter = "VAN_ALBADA"
venkat_k = 1.5
gradient = "LEAST_SQUARES_FACE"
[time]
mode = "Steady"
scheme = "BACKWARD_EULER"
cfl = 20
max_iterations = 100
residual_tolerance = 1e-9
[output]
directory = "o"
field_interval = 0
residual_interval = 10
[linalg]
preconditioner = "SGS"
rel_tol = 0.05
max_iter = 60
"""

BC_TYPES = [
    'type = "SUPERSONIC_INLET"\np = 9e4\nt = 280\nvelocity = [400, 0, 10]\n',
    'type = "SUBSONIC_OUTLET"\np_back = 95000\n',
    'type = "NO_SLIP_WALL"\nt_wall = 310\n',
    'type = "FARFIELD"\nmach = 0.6\np_inf = 1e5\nt_inf = 290\nalpha_deg = 3\n',
    'type = "SYMMETRY"\n',
    'type = "NO_SLIP_WALL_HEAT_FLUX"\ntmp_grad = 2.5\n',
    'type = "SUBSONIC_INLET"\nt = 300\nmach = 0.2\ndirection = [1, 1, 0]\n',
]


def integration_blocks():
    text = open(os.path.join(ROOT, "INTEGRATION.md"), encoding="utf-8").read()
    s1 = text.index("## 1. Mesh")
    s3 = text.index("## 3. The run loop")
    blocks = re.findall(r"```cpp\n(.*?)```", text[s1:s3], flags=re.S)
    assert len(blocks) == 2, "sections 1 and 2 each hold one C++ block"
    return blocks


def test_integration_snippets_compile_and_bind(tmp_path):
    blocks = integration_blocks()
    src = tmp_path / "binding.cpp"
    src.write_text("\n".join(blocks) + MAIN, encoding="utf-8")
    exe = tmp_path / "binding"
    lib_dir = os.path.dirname(cfdg.LIB_PATH)
    gxx = shutil.which("g++") or "g++"
    env = dict(os.environ, MPICH_CXX=gxx)
    obj = tmp_path / "binding.o"
    cmd = [MPICXX, "-std=c++20", "-O1", "-DNDEBUG", "-Wall", f"-I{REF}/include",
           f"-I{REF}/external/tomlplusplus/include", f"-I{ROOT}/include", "-c", str(src), "-o", str(obj)]
    p = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600)
    assert p.returncode == 0, p.stderr[-4000:]
    # linked with the system compiler and its static libstdc++, as oracle/build_ref.sh links the
    # reference's objects (the MPI wrapper's library path holds an older libstdc++); libcfdg.so keeps
    # its own, newer libstdc++ (RUNPATH entries apply to direct dependencies only), so its versioned
    # C++ runtime references are left for the loader
    cmd = [gxx, str(obj)] + [os.path.join(OBJ, o) for o in REF_OBJS]
    cmd += [f"-L{lib_dir}", "-lcfdg", "/opt/conda/lib/libmpi.so", "-Wl,--enable-new-dtags",
            f"-Wl,-rpath,{lib_dir}", "-Wl,-rpath,/opt/conda/lib", "-Wl,--allow-shlib-undefined",
            "-static-libstdc++", "-static-libgcc", "-o", str(exe)]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-4000:]
    (tmp_path / "solver.toml").write_text(SOLVER, encoding="utf-8")
    m = cfdg.Mesh.generate(kind=1, ni=4, nj=3, nk=2, x1=2.0, y1=1.0, z1=0.5, bump_h=0.05, bump_w=0.3, bump_x=1.0,
                           seed=5, reorder=1)
    bc = "\n".join(f'[[boundary_condition]]\npatch_id = {p}\n' + BC_TYPES[p % len(BC_TYPES)]
                   for p in range(m.n_patches))
    (tmp_path / "bc.toml").write_text(bc, encoding="utf-8")
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "binding ok" in r.stdout
