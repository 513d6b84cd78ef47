// Driver that runs the REFERENCE solver's own compiled code on a mesh and
// state given in a record file (oracle/recio.py) and writes every
// intermediate array back.  Built by oracle/build_ref.sh from the sources under
// the reference tree into oracle/_ref/ (never committed).  It only calls the
// reference's public classes and functions; it contains no physics of its own.
//
// usage: ref_harness <in.rec> <out.rec>
#include <mpi.h>

#include <array>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

#include "cfd/bc/config.hpp"
#include "cfd/core/types.hpp"
#include "cfd/linalg/config.hpp"
#include "cfd/mesh/aux_connectivity.hpp"
#include "cfd/mesh/aux_geometry.hpp"
#include "cfd/mesh/geometry.hpp"
#include "cfd/mesh/localmesh.hpp"
#include "cfd/mesh/reorder.hpp"
#include "cfd/mesh/validate.hpp"
#include "cfd/mpi/log.hpp"
#include "cfd/numerics/gradient/gradient_manager.hpp"
#include "cfd/numerics/limiter/slope_limiters.hpp"
#include "cfd/numerics/reconstruction/first_order.hpp"
#include "cfd/numerics/reconstruction/muscl.hpp"
#include "cfd/solver/bsr_linear_system.hpp"
#include "cfd/solver/compressible/bc_manager.hpp"
#include "cfd/solver/compressible/config.hpp"
#include "cfd/solver/compressible/eos/ideal_gas.hpp"
#include "cfd/solver/compressible/eos/state_conversions.hpp"
#include "cfd/solver/compressible/fluxes/hllc.hpp"
#include "cfd/solver/compressible/jacobian_kernel.hpp"
#include "cfd/solver/compressible/physics/inviscid_flow.hpp"
#include "cfd/solver/compressible/physics/stack.hpp"
#include "cfd/solver/compressible/physics/viscous_flow.hpp"
#include "cfd/solver/compressible/residual_kernel.hpp"
#include "cfd/solver/compressible/solver.hpp"
#include "cfd/solver/compressible/time/forward_euler.hpp"
#include "cfd/solver/compressible/time/implicit_euler.hpp"
#include "cfd/solver/time_mode.hpp"

namespace {

using namespace cfd;
namespace sc = cfd::solver::compressible;

struct Rec {
    int code;
    std::vector<char> data;
    std::size_t count;
};
using Recs = std::map<std::string, Rec>;

Recs read_recs(const char* path) {
    std::ifstream f(path, std::ios::binary);
    char magic[8];
    f.read(magic, 8);
    if (!f || std::memcmp(magic, "CFDGREC1", 8) != 0) throw std::runtime_error("bad record file");
    Recs r;
    while (true) {
        std::uint32_t ln = 0;
        if (!f.read(reinterpret_cast<char*>(&ln), 4)) break;
        std::string name(ln, '\0');
        f.read(name.data(), ln);
        unsigned char code = 0;
        std::uint64_t count = 0;
        f.read(reinterpret_cast<char*>(&code), 1);
        f.read(reinterpret_cast<char*>(&count), 8);
        const std::size_t es = (code == 0 || code == 3) ? 8 : (code == 1 ? 4 : 1);
        Rec rc{code, std::vector<char>(count * es), count};
        f.read(rc.data.data(), static_cast<std::streamsize>(count * es));
        r[name] = std::move(rc);
    }
    return r;
}

template <typename T>
std::vector<T> get(const Recs& r, const std::string& n) {
    const auto it = r.find(n);
    if (it == r.end()) throw std::runtime_error("missing record " + n);
    std::vector<T> v(it->second.count);
    std::memcpy(v.data(), it->second.data.data(), it->second.data.size());
    return v;
}

template <typename T>
T scalar(const Recs& r, const std::string& n) {
    return get<T>(r, n).at(0);
}

struct Writer {
    std::ofstream f;
    explicit Writer(const char* p) : f(p, std::ios::binary) { f.write("CFDGREC1", 8); }
    template <typename T>
    void put(const std::string& n, const T* d, std::size_t count) {
        const std::uint32_t ln = static_cast<std::uint32_t>(n.size());
        f.write(reinterpret_cast<const char*>(&ln), 4);
        f.write(n.data(), ln);
        unsigned char code = std::is_same_v<T, double>         ? 0
                             : std::is_same_v<T, std::int32_t> ? 1
                             : std::is_same_v<T, std::uint8_t> ? 2
                                                               : 3;
        const std::uint64_t c = count;
        f.write(reinterpret_cast<const char*>(&code), 1);
        f.write(reinterpret_cast<const char*>(&c), 8);
        f.write(reinterpret_cast<const char*>(d), static_cast<std::streamsize>(count * sizeof(T)));
    }
    template <typename T>
    void put(const std::string& n, const std::vector<T>& v) {
        put(n, v.data(), v.size());
    }
};

mesh::MeshPart build_meshpart(const Recs& r) {
    mesh::MeshPart mp;
    mp.rank = 0;
    mp.nprocs = 1;
    const auto vol = get<double>(r, "cell_vol");
    const auto ctype = get<std::uint8_t>(r, "cell_type");
    const auto nx = get<double>(r, "node_x");
    const auto owner = get<std::int32_t>(r, "face_owner");
    const auto neigh = get<std::int32_t>(r, "face_neigh");
    const LocalIndex nc = static_cast<LocalIndex>(vol.size());
    const LocalIndex nf = static_cast<LocalIndex>(owner.size());
    mp.n_cells_g = nc;
    mp.n_nodes_g = static_cast<GlobalIndex>(nx.size());
    mp.n_faces_g = nf;
    mp.n_own = nc;
    mp.n_cells = nc;
    mp.cell_type.resize(static_cast<std::size_t>(nc));
    for (LocalIndex c = 0; c < nc; ++c) mp.cell_type[static_cast<std::size_t>(c)] = static_cast<mesh::CellType>(ctype[static_cast<std::size_t>(c)]);
    mp.cell_gid.resize(static_cast<std::size_t>(nc));
    for (LocalIndex c = 0; c < nc; ++c) mp.cell_gid[static_cast<std::size_t>(c)] = c;
    mp.cell_donor.assign(static_cast<std::size_t>(nc), -1);
    mp.cell_nodes_offsets = get<std::int32_t>(r, "cell_nodes_off");
    mp.cell_nodes = get<std::int32_t>(r, "cell_nodes");
    mp.cell_centroid_x = get<double>(r, "cell_cx");
    mp.cell_centroid_y = get<double>(r, "cell_cy");
    mp.cell_centroid_z = get<double>(r, "cell_cz");
    mp.cell_volume = vol;
    mp.n_nodes = static_cast<LocalIndex>(nx.size());
    mp.n_nodes_own = mp.n_nodes;
    mp.node_gid.resize(nx.size());
    for (std::size_t i = 0; i < nx.size(); ++i) mp.node_gid[i] = static_cast<GlobalIndex>(i);
    mp.node_x = nx;
    mp.node_y = get<double>(r, "node_y");
    mp.node_z = get<double>(r, "node_z");
    mp.n_faces = nf;
    LocalIndex ninner = 0;
    for (const auto n : neigh)
        if (n >= 0) ++ninner;
    mp.n_inner_faces = ninner;
    mp.n_internal_faces = ninner;
    mp.face_owner = owner;
    mp.face_neigh = neigh;
    const auto ftype = get<std::uint8_t>(r, "face_type");
    mp.face_type.resize(ftype.size());
    for (std::size_t i = 0; i < ftype.size(); ++i) mp.face_type[i] = static_cast<mesh::CellType>(ftype[i]);
    mp.face_nodes_offsets = get<std::int32_t>(r, "face_nodes_off");
    mp.face_nodes = get<std::int32_t>(r, "face_nodes");
    mp.face_centroid_x = get<double>(r, "face_cx");
    mp.face_centroid_y = get<double>(r, "face_cy");
    mp.face_centroid_z = get<double>(r, "face_cz");
    mp.face_normal_x = get<double>(r, "face_nx");
    mp.face_normal_y = get<double>(r, "face_ny");
    mp.face_normal_z = get<double>(r, "face_nz");
    mp.face_area = get<double>(r, "face_area");
    mp.face_patch = get<std::int32_t>(r, "face_patch");
    mp.n_bfaces_g = nf - ninner;
    const int np = scalar<std::int32_t>(r, "n_patches");
    for (int p = 0; p < np; ++p) mp.patches.push_back({"patch" + std::to_string(p), "UserDefined"});
    mp.recv_offsets.assign(1, 0);
    mp.send_offsets.assign(1, 0);
    mp.patch_face_offsets.assign(static_cast<std::size_t>(np) + 1, 0);
    for (LocalIndex f = ninner; f < nf; ++f) ++mp.patch_face_offsets[static_cast<std::size_t>(mp.face_patch[static_cast<std::size_t>(f)]) + 1];
    for (int p = 0; p < np; ++p) mp.patch_face_offsets[static_cast<std::size_t>(p) + 1] += mp.patch_face_offsets[static_cast<std::size_t>(p)];
    for (LocalIndex f = ninner; f < nf; ++f) mp.patch_faces.push_back(f);
    for (int d = 0; d < 3; ++d) {
        mp.bbox_lo[d] = 0.0;
        mp.bbox_hi[d] = 0.0;
    }
    return mp;
}

bc::BoundaryConfig build_bcs(const Recs& r) {
    const int np = scalar<std::int32_t>(r, "n_patches");
    const auto type = get<std::int32_t>(r, "bc_type");
    const auto mode = get<std::int32_t>(r, "bc_inflow_mode");
    const auto sc_ = get<double>(r, "bc_scalars");  // p, t, tmp_grad, mach, alpha, beta per patch
    const auto vel = get<double>(r, "bc_velocity");
    const auto dir = get<double>(r, "bc_direction");
    bc::BoundaryConfig out;
    out.patches.resize(static_cast<std::size_t>(np));
    for (int p = 0; p < np; ++p) {
        const auto ps = static_cast<std::size_t>(p);
        bc::BCDescriptor& d = out.patches[ps];
        d.patch_id = p;
        d.type = static_cast<bc::physical::BCType>(type[ps]);
        d.inflow_mode = static_cast<bc::physical::InflowMode>(mode[ps]);
        d.p = sc_[6 * ps + 0];
        d.t = sc_[6 * ps + 1];
        d.tmp_grad = sc_[6 * ps + 2];
        d.mach = sc_[6 * ps + 3];
        d.alpha_deg = sc_[6 * ps + 4];
        d.beta_deg = sc_[6 * ps + 5];
        for (int k = 0; k < 3; ++k) {
            d.velocity[static_cast<std::size_t>(k)] = vel[3 * ps + static_cast<std::size_t>(k)];
            d.direction[static_cast<std::size_t>(k)] = dir[3 * ps + static_cast<std::size_t>(k)];
        }
    }
    return out;
}

numerics::gradient::GradientType grad_type(int g) {
    if (g == 0) return numerics::gradient::GradientType::GreenGaussFace;
    if (g == 1) return numerics::gradient::GradientType::GreenGaussCell;
    return numerics::gradient::GradientType::LeastSquaresCellNode;
}

// Residual pipeline assembled from the reference's public components, in the
// order of Solver::evaluate_residual (solver.hpp:169-247).
template <typename Phys, typename Recon>
void run_components(const Recs& r, const mesh::MeshPart& mp, const bc::BoundaryConfig& bcfg, const Phys& phys,
                    Writer& w) {
    using EOS = sc::eos::IdealGas;
    const EOS eos{scalar<double>(r, "gamma"), scalar<double>(r, "gas_constant")};
    const std::size_t nc = static_cast<std::size_t>(mp.n_cells);
    const std::size_t nt = nc + static_cast<std::size_t>(mp.n_faces - mp.n_inner_faces);
    const auto U = get<double>(r, "U");  // 5 x nc

    mesh::MeshAuxConnectivity aux_conn;
    aux_conn.add_connectivity(mp, mesh::AuxConnType::CellFaces | mesh::AuxConnType::CellCellsByFace);
    mesh::MeshAuxGeometry aux_geom;
    aux_geom.add_geometry(mp, Recon::kAuxGeometry | Phys::kAuxGeometry);

    std::vector<std::vector<double>> q(5, std::vector<double>(nt, 0.0)), gx(5, std::vector<double>(nt, 0.0)),
        gy(5, std::vector<double>(nt, 0.0)), gz(5, std::vector<double>(nt, 0.0)), phi(5, std::vector<double>(nt, 0.0)),
        res(5, std::vector<double>(nt, 0.0));
    std::vector<double> lam(nc, 0.0);
    std::vector<double*> qp, gxp, gyp, gzp, php, rp;
    for (int v = 0; v < 5; ++v) {
        qp.push_back(q[v].data());
        gxp.push_back(gx[v].data());
        gyp.push_back(gy[v].data());
        gzp.push_back(gz[v].data());
        php.push_back(phi[v].data());
        rp.push_back(res[v].data());
    }
    const std::vector<const double*> qc(qp.begin(), qp.end()), gxc(gxp.begin(), gxp.end()), gyc(gyp.begin(), gyp.end()),
        gzc(gzp.begin(), gzp.end()), phc(php.begin(), php.end());

    for (std::size_t c = 0; c < nc; ++c) {
        const double Uc[5] = {U[c], U[nc + c], U[2 * nc + c], U[3 * nc + c], U[4 * nc + c]};
        sc::eos::conserved_to_primitives_pT(eos, Uc, q[0][c], q[1][c], q[2][c], q[3][c], q[4][c]);
    }
    sc::BoundaryManager<EOS> bcs;
    bcs.initialize(bcfg, mp, eos);
    bcs.update_ghost_cells(qp, mp);
    const bool need_grad = Recon::kNeedsGradients || Phys::kNeedsGradients;
    if (need_grad) {
        numerics::gradient::GradientManager gm;
        gm.create_gradient(grad_type(scalar<std::int32_t>(r, "gradient")));
        gm.setup_gradient(mp, aux_conn);
        gm.apply_gradient_set(qc, gxp, gyp, gzp, mp, aux_conn);
        bcs.update_ghost_cells_grad(qc, gxp, gyp, gzp, mp);
        if constexpr (Recon::kNeedsGradients) {
            Recon::template compute_limiters<5>(mp, aux_conn, qc.data(), gxc.data(), gyc.data(), gzc.data(), php.data(),
                                                scalar<double>(r, "venkat_k"));
        }
    }
    // an eddy-viscosity field given as input goes to the kernels' mut argument
    // (residual_kernel.hpp:64-91, 96-169; jacobian_kernel.hpp:74-90)
    const std::vector<double> mut = r.count("mut") ? get<double>(r, "mut") : std::vector<double>{};
    if (!mut.empty() && mut.size() != nt) throw std::runtime_error("mut needs n_cells + boundary faces values");
    const double* mutp = mut.empty() ? nullptr : mut.data();
    sc::ResidualKernel<EOS, sc::fluxes::HllcFlux, Recon, Phys> rk(mp, aux_conn, aux_geom, eos, phys);
    rk.apply(qc, gxc, gyc, gzc, phc, rp, lam.data(), mutp);
    std::vector<double> bnd_mass, bnd_energy;
    rk.boundary_integrals(qc, gxc, gyc, gzc, phc, bnd_mass, bnd_energy, mutp);
    if (r.count("jac_alpha")) {
        // JacobianKernel::apply on these primitives, the given diagonal shift
        // and the eddy-viscosity field, into the reference's own BSR pattern
        const auto alpha = get<double>(r, "jac_alpha");
        if (alpha.size() != nc) throw std::runtime_error("jac_alpha needs n_cells values");
        linalg::SolverParams lpar;
        solver::BSRLinearSystem<5> sys(mp, aux_conn, lpar, MPI_COMM_WORLD);
        sc::JacobianKernel<EOS, sc::fluxes::HllcFlux, Phys> jk(mp, aux_conn, aux_geom, eos, phys);
        jk.apply(qc, sys.matrix().rowPtr().data(), sys.matrix().cols().data(), sys.matrix().diagIndex().data(),
                 sys.matrix().valuesData(), alpha.data(), mutp);
        w.put("jac_values", sys.matrix().values());
    }

    std::vector<double> qo, go, po, ro;
    for (int v = 0; v < 5; ++v) qo.insert(qo.end(), q[v].begin(), q[v].end());
    for (int v = 0; v < 5; ++v) {
        go.insert(go.end(), gx[v].begin(), gx[v].end());
        go.insert(go.end(), gy[v].begin(), gy[v].end());
        go.insert(go.end(), gz[v].begin(), gz[v].end());
    }
    for (int v = 0; v < 5; ++v) po.insert(po.end(), phi[v].begin(), phi[v].begin() + static_cast<std::ptrdiff_t>(nc));
    for (int v = 0; v < 5; ++v) ro.insert(ro.end(), res[v].begin(), res[v].begin() + static_cast<std::ptrdiff_t>(nc));
    w.put("q", qo);
    w.put("grad", go);
    w.put("phi", po);
    w.put("res", ro);
    w.put("lam", lam);
    w.put("bnd_mass", bnd_mass);
    w.put("bnd_energy", bnd_energy);
    // wall distance from the no-slip patches (MeshAuxGeometry::add_wall_distance), both stencils
    {
        const auto types = get<std::int32_t>(r, "bc_type");
        std::vector<bool> wall(mp.patches.size(), false);
        for (std::size_t p = 0; p < wall.size() && p < types.size(); ++p)
            wall[p] = types[p] == 5 || types[p] == 6;  // NO_SLIP_WALL, NO_SLIP_WALL_HEAT_FLUX
        mesh::MeshAuxConnectivity wc;
        mesh::MeshAuxGeometry wf, wn;
        wf.add_wall_distance(mp, wc, MPI_COMM_WORLD, wall, mesh::WallDistStencil::FaceNeighbors, 100, 1.0e-4);
        wn.add_wall_distance(mp, wc, MPI_COMM_WORLD, wall, mesh::WallDistStencil::NodeNeighbors, 500, 1.0e-8);
        w.put("wall_dist_face", wf.cell_wall_dist);
        w.put("wall_dist_node", wn.cell_wall_dist);
    }
    w.put("ccn_off", aux_conn.cell_cells_face_offsets);  // face stencil of the limiter
    if (need_grad && scalar<std::int32_t>(r, "gradient") == 2) {
        w.put("lsq_off", aux_conn.cell_cells_node_offsets);
        w.put("lsq_nb", aux_conn.cell_cells_node);
    }
}

// Full Solver object: evaluate_residual, compute_dt, assemble_jacobian into a
// BSRLinearSystem<5>, one linear solve; optionally N implicit steps via run().
template <typename Phys, typename Recon>
void run_solver_path(const Recs& r, const mesh::MeshPart& mp, const bc::BoundaryConfig& bcfg, const Phys& phys,
                     Writer& w) {
    using EOS = sc::eos::IdealGas;
    const EOS eos{scalar<double>(r, "gamma"), scalar<double>(r, "gas_constant")};
    sc::SolverConfig cfg;
    cfg.flow.gamma = eos.gamma_val;
    cfg.flow.gas_constant = eos.R_val;
    cfg.prandtl = scalar<double>(r, "prandtl");
    cfg.flow_model = Phys::kHasViscous ? sc::FlowModel::ViscousFlow : sc::FlowModel::InviscidFlow;
    const auto init = get<double>(r, "init");  // rho, u, v, w, p
    cfg.init_rho = init[0];
    cfg.init_velocity = {init[1], init[2], init[3]};
    cfg.init_p = init[4];
    cfg.gradient = grad_type(scalar<std::int32_t>(r, "gradient"));
    cfg.limiter_venkat_k = scalar<double>(r, "venkat_k");
    cfg.scheme = sc::TimeScheme::BackwardEuler;
    cfg.cfl = scalar<double>(r, "cfl");
    const auto steps = scalar<std::int32_t>(r, "steps");
    cfg.max_iterations = steps > 0 ? steps : 1;
    cfg.residual_tolerance = 1e-300;
    cfg.field_interval = 0;
    cfg.residual_interval = 1;
    cfg.output_dir = "/tmp/cfdg_ref_out";
    cfg.linear_solver_params.type = linalg::SolverType::BICGSTAB;
    cfg.linear_solver_params.precond_type =
        scalar<std::int32_t>(r, "precond") == 1 ? linalg::PreconditionerType::SGS : linalg::PreconditionerType::None;
    cfg.linear_solver_params.relative_tolerance = scalar<double>(r, "rel_tol");
    cfg.linear_solver_params.absolute_tolerance = scalar<double>(r, "abs_tol");
    cfg.linear_solver_params.max_iterations = scalar<std::int32_t>(r, "max_iter");

    using SolverT = sc::Solver<EOS, sc::fluxes::HllcFlux, Recon, Phys, sc::time::BackwardEuler, solver::SteadyMode>;
    SolverT s(cfg, bcfg, eos, phys, mp, MPI_COMM_WORLD);

    const std::size_t nc = static_cast<std::size_t>(mp.n_cells);
    const auto U = get<double>(r, "U");
    auto us = s.u_slots();
    for (int v = 0; v < 5; ++v)
        for (std::size_t c = 0; c < nc; ++c) us[static_cast<std::size_t>(v)][c] = U[static_cast<std::size_t>(v) * nc + c];
    s.evaluate_residual(s.u_slots());
    s.compute_dt();

    mesh::MeshAuxConnectivity aux_conn;
    aux_conn.add_connectivity(mp, mesh::AuxConnType::CellCellsByFace);
    solver::BSRLinearSystem<5> sys(mp, aux_conn, cfg.linear_solver_params, MPI_COMM_WORLD);
    s.assemble_jacobian(sys.matrix().rowPtr().data(), sys.matrix().cols().data(), sys.matrix().diagIndex().data(),
                        sys.matrix().valuesData());
    {
        std::vector<double> ro;
        const auto rs = s.res_slots();
        for (int v = 0; v < 5; ++v) ro.insert(ro.end(), rs[static_cast<std::size_t>(v)], rs[static_cast<std::size_t>(v)] + nc);
        w.put("solver_res", ro);
        w.put("dt", s.local_dt(), nc);
        w.put("alpha", s.alpha(), nc);
        w.put("row_ptr", sys.matrix().rowPtr());
        w.put("cols", sys.matrix().cols());
        w.put("diag", sys.matrix().diagIndex());
        w.put("values", sys.matrix().values());
        double* b = sys.rhs_data();
        double* du = sys.du_data();
        for (std::size_t c = 0; c < nc; ++c)
            for (int v = 0; v < 5; ++v) {
                b[c * 5 + static_cast<std::size_t>(v)] = -rs[static_cast<std::size_t>(v)][c];
                du[c * 5 + static_cast<std::size_t>(v)] = 0.0;
            }
        sys.solve();
        w.put("du", du, 5 * nc);
        const auto& lr = sys.last_result();
        const std::int32_t st[2] = {static_cast<std::int32_t>(lr.status), lr.iterations};
        w.put("lin_status", st, 2);
        const double nr[2] = {lr.initial_residual, lr.final_residual};
        w.put("lin_norms", nr, 2);
    }
    if (steps > 0) {
        SolverT s2(cfg, bcfg, eos, phys, mp, MPI_COMM_WORLD);
        s2.run();
        std::vector<double> uo;
        const auto u2 = s2.u_slots();
        for (int v = 0; v < 5; ++v) uo.insert(uo.end(), u2[static_cast<std::size_t>(v)], u2[static_cast<std::size_t>(v)] + nc);
        w.put("steps_U", uo);
        // the solution files run() wrote at the end (Solver::write_fields, vtu.cpp:321-424)
        char stem[96];
        std::snprintf(stem, sizeof(stem), "%s/final_%09lld", cfg.output_dir.c_str(), static_cast<long long>(steps));
        const auto slurp = [](const std::string& path) {
            std::ifstream in(path, std::ios::binary);
            return std::vector<std::uint8_t>((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        };
        w.put("vtu_piece", slurp(std::string(stem) + "_00000.vtu"));
        w.put("vtu_master", slurp(std::string(stem) + ".pvtu"));
    }
}

// Time stepping through Solver::run from the uniform initial state: N forward
// Euler iterations (forward_euler.hpp) and/or the dual-time BDF1/BDF2 driver
// (solver.hpp:662-736, BackwardEuler inner iterations).
template <typename Phys, typename Recon>
void run_time_paths(const Recs& r, const mesh::MeshPart& mp, const bc::BoundaryConfig& bcfg, const Phys& phys,
                    Writer& w) {
    using EOS = sc::eos::IdealGas;
    const EOS eos{scalar<double>(r, "gamma"), scalar<double>(r, "gas_constant")};
    sc::SolverConfig cfg;
    cfg.flow.gamma = eos.gamma_val;
    cfg.flow.gas_constant = eos.R_val;
    cfg.prandtl = scalar<double>(r, "prandtl");
    cfg.flow_model = Phys::kHasViscous ? sc::FlowModel::ViscousFlow : sc::FlowModel::InviscidFlow;
    const auto init = get<double>(r, "init");  // rho, u, v, w, p
    cfg.init_rho = init[0];
    cfg.init_velocity = {init[1], init[2], init[3]};
    cfg.init_p = init[4];
    cfg.gradient = grad_type(scalar<std::int32_t>(r, "gradient"));
    cfg.limiter_venkat_k = scalar<double>(r, "venkat_k");
    cfg.field_interval = 0;
    cfg.residual_interval = 1;
    cfg.output_dir = "/tmp/cfdg_ref_out";
    cfg.linear_solver_params.type = linalg::SolverType::BICGSTAB;
    cfg.linear_solver_params.precond_type =
        scalar<std::int32_t>(r, "precond") == 1 ? linalg::PreconditionerType::SGS : linalg::PreconditionerType::None;
    cfg.linear_solver_params.relative_tolerance = scalar<double>(r, "rel_tol");
    cfg.linear_solver_params.absolute_tolerance = scalar<double>(r, "abs_tol");
    cfg.linear_solver_params.max_iterations = scalar<std::int32_t>(r, "max_iter");
    const std::size_t nc = static_cast<std::size_t>(mp.n_cells);
    const auto dump = [&](const char* name, auto& solver) {
        std::vector<double> uo;
        const auto u = solver.u_slots();
        for (int v = 0; v < 5; ++v) uo.insert(uo.end(), u[static_cast<std::size_t>(v)], u[static_cast<std::size_t>(v)] + nc);
        w.put(name, uo);
    };
    if (r.count("explicit_steps") && scalar<std::int32_t>(r, "explicit_steps") > 0) {
        sc::SolverConfig c = cfg;
        c.scheme = sc::TimeScheme::ForwardEuler;
        c.cfl = scalar<double>(r, "explicit_cfl");
        c.max_iterations = scalar<std::int32_t>(r, "explicit_steps");
        c.residual_tolerance = 1e-300;
        sc::Solver<EOS, sc::fluxes::HllcFlux, Recon, Phys, sc::time::ForwardEuler, solver::SteadyMode> s(
            c, bcfg, eos, phys, mp, MPI_COMM_WORLD);
        s.run();
        dump("explicit_U", s);
    }
    if (r.count("unsteady_steps") && scalar<std::int32_t>(r, "unsteady_steps") > 0) {
        sc::SolverConfig c = cfg;
        c.scheme = sc::TimeScheme::BackwardEuler;
        c.cfl = scalar<double>(r, "cfl");
        c.dt = scalar<double>(r, "unsteady_dt");
        c.max_time_steps = scalar<std::int32_t>(r, "unsteady_steps");
        c.max_iterations = scalar<std::int32_t>(r, "unsteady_subiter");
        c.residual_tolerance = scalar<double>(r, "unsteady_sub_tol");
        if (scalar<std::int32_t>(r, "unsteady_order") == 1) {
            sc::Solver<EOS, sc::fluxes::HllcFlux, Recon, Phys, sc::time::BackwardEuler, solver::UnsteadyMode<1>> s(
                c, bcfg, eos, phys, mp, MPI_COMM_WORLD);
            s.run();
            dump("unsteady_U", s);
        } else {
            sc::Solver<EOS, sc::fluxes::HllcFlux, Recon, Phys, sc::time::BackwardEuler, solver::UnsteadyMode<2>> s(
                c, bcfg, eos, phys, mp, MPI_COMM_WORLD);
            s.run();
            dump("unsteady_U", s);
        }
    }
}

// Wall time of Solver::evaluate_residual on the given state (MPI_Wtime around
// `reps` calls), the reference's own CPU cost of one residual evaluation.
template <typename Phys, typename Recon>
void run_timing(const Recs& r, const mesh::MeshPart& mp, const bc::BoundaryConfig& bcfg, const Phys& phys, Writer& w) {
    using EOS = sc::eos::IdealGas;
    const EOS eos{scalar<double>(r, "gamma"), scalar<double>(r, "gas_constant")};
    sc::SolverConfig cfg;
    cfg.flow.gamma = eos.gamma_val;
    cfg.flow.gas_constant = eos.R_val;
    cfg.prandtl = scalar<double>(r, "prandtl");
    cfg.flow_model = Phys::kHasViscous ? sc::FlowModel::ViscousFlow : sc::FlowModel::InviscidFlow;
    cfg.gradient = grad_type(scalar<std::int32_t>(r, "gradient"));
    cfg.limiter_venkat_k = scalar<double>(r, "venkat_k");
    cfg.scheme = sc::TimeScheme::BackwardEuler;
    cfg.cfl = scalar<double>(r, "cfl");
    cfg.output_dir = "/tmp/cfdg_ref_out";
    const double t0 = MPI_Wtime();
    sc::Solver<EOS, sc::fluxes::HllcFlux, Recon, Phys, sc::time::BackwardEuler, solver::SteadyMode> s(
        cfg, bcfg, eos, phys, mp, MPI_COMM_WORLD);
    const double t_setup = MPI_Wtime() - t0;
    const std::size_t nc = static_cast<std::size_t>(mp.n_cells);
    const auto U = get<double>(r, "U");
    auto us = s.u_slots();
    for (int v = 0; v < 5; ++v)
        for (std::size_t c = 0; c < nc; ++c) us[static_cast<std::size_t>(v)][c] = U[static_cast<std::size_t>(v) * nc + c];
    const int reps = scalar<std::int32_t>(r, "time_reps");
    std::vector<double> times;
    for (int i = 0; i < reps; ++i) {
        const double a = MPI_Wtime();
        s.evaluate_residual(s.u_slots());
        times.push_back(MPI_Wtime() - a);
    }
    w.put("time_residual", times);
    w.put("time_setup", &t_setup, 1);
}

// BSRLinearSystem<5>::solve (BiCGSTAB + the configured preconditioner,
// bsr_linear_system.hpp:93-97) on given block values, right-hand side and
// initial guess, in the reference's own pattern of the mesh.
void run_linear_system(const Recs& r, const mesh::MeshPart& mp, Writer& w) {
    mesh::MeshAuxConnectivity aux_conn;
    aux_conn.add_connectivity(mp, mesh::AuxConnType::CellCellsByFace);
    linalg::SolverParams lp;
    lp.type = linalg::SolverType::BICGSTAB;
    lp.precond_type =
        scalar<std::int32_t>(r, "precond") == 1 ? linalg::PreconditionerType::SGS : linalg::PreconditionerType::None;
    lp.relative_tolerance = scalar<double>(r, "rel_tol");
    lp.absolute_tolerance = scalar<double>(r, "abs_tol");
    lp.max_iterations = scalar<std::int32_t>(r, "max_iter");
    solver::BSRLinearSystem<5> sys(mp, aux_conn, lp, MPI_COMM_WORLD);
    const auto vals = get<double>(r, "lin_values");
    const auto b = get<double>(r, "lin_b");
    const auto x0 = get<double>(r, "lin_x0");
    const std::size_t n5 = 5 * static_cast<std::size_t>(mp.n_cells);
    if (vals.size() != sys.matrix().values().size() || b.size() != n5 || x0.size() != n5)
        throw std::runtime_error("lin_values / lin_b / lin_x0 sizes");
    std::copy(vals.begin(), vals.end(), sys.matrix().valuesData());
    std::copy(b.begin(), b.end(), sys.rhs_data());
    std::copy(x0.begin(), x0.end(), sys.du_data());
    sys.solve();
    w.put("lin_x", sys.du_data(), n5);
    const auto& lr = sys.last_result();
    const std::int32_t st[2] = {static_cast<std::int32_t>(lr.status), lr.iterations};
    w.put("lin_status", st, 2);
    const double nr[2] = {lr.initial_residual, lr.final_residual};
    w.put("lin_norms", nr, 2);
    w.put("row_ptr", sys.matrix().rowPtr());
    w.put("cols", sys.matrix().cols());
    w.put("diag", sys.matrix().diagIndex());
}

template <typename Phys, typename Recon>
void run_all(const Recs& r, const mesh::MeshPart& mp, const bc::BoundaryConfig& bcfg, const Phys& phys, Writer& w) {
    if (r.count("lin_values")) {
        run_linear_system(r, mp, w);
        return;
    }
    if (r.count("time_reps") && scalar<std::int32_t>(r, "time_reps") > 0) {
        run_timing<Phys, Recon>(r, mp, bcfg, phys, w);
        return;
    }
    run_components<Phys, Recon>(r, mp, bcfg, phys, w);
    if (scalar<std::int32_t>(r, "solver_path") != 0) run_solver_path<Phys, Recon>(r, mp, bcfg, phys, w);
    run_time_paths<Phys, Recon>(r, mp, bcfg, phys, w);
}

template <typename Phys>
void dispatch_recon(const Recs& r, const mesh::MeshPart& mp, const bc::BoundaryConfig& bcfg, const Phys& phys, Writer& w) {
    const int recon = scalar<std::int32_t>(r, "reconstruction");
    const int lim = scalar<std::int32_t>(r, "limiter");
    using namespace numerics;
    if (recon == 0) return run_all<Phys, recon::FirstOrder>(r, mp, bcfg, phys, w);
    if (lim == 0) return run_all<Phys, recon::Muscl<limiter::Venkatakrishnan>>(r, mp, bcfg, phys, w);
    if (lim == 1) return run_all<Phys, recon::Muscl<limiter::BarthJespersen>>(r, mp, bcfg, phys, w);
    if (lim == 2) return run_all<Phys, recon::Muscl<limiter::VanAlbada>>(r, mp, bcfg, phys, w);
    return run_all<Phys, recon::Muscl<limiter::Unlimited>>(r, mp, bcfg, phys, w);
}

}  // namespace

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    int rc = 0;
    try {
        if (argc != 3) throw std::runtime_error("usage: ref_harness <in.rec> <out.rec>");
        mpi::log_init(0);
        const Recs r = read_recs(argv[1]);
        mesh::MeshPart mp = build_meshpart(r);
        Writer w(argv[2]);
        if (scalar<std::int32_t>(r, "validate") != 0) mesh::validate_and_log_meshpart(mp);
        if (r.count("rcm") && scalar<std::int32_t>(r, "rcm") != 0) {
            // reference renumbering of the given mesh (reorder.cpp:579-602)
            mesh::MeshPart m2 = mp;
            mesh::reorder_local_mesh(m2, mesh::ReorderMethod::RCM);
            std::vector<std::int64_t> gid(m2.cell_gid.begin(), m2.cell_gid.end());
            w.put("rcm_new2old", gid);
            w.put("rcm_face_owner", m2.face_owner);
            w.put("rcm_face_neigh", m2.face_neigh);
            w.put("rcm_face_patch", m2.face_patch);
            w.put("rcm_face_area", m2.face_area);
            w.put("rcm_cell_cx", m2.cell_centroid_x);
        }
        if (r.count("geometry") && scalar<std::int32_t>(r, "geometry") != 0) {
            // reference cell and face metrics from the nodes and connectivity
            // alone (geometry.cpp:301-539); the given metrics are discarded
            mesh::MeshPart g = mp;
            const double nan = std::numeric_limits<double>::quiet_NaN();
            for (auto* v : {&g.cell_centroid_x, &g.cell_centroid_y, &g.cell_centroid_z, &g.cell_volume,
                            &g.face_centroid_x, &g.face_centroid_y, &g.face_centroid_z, &g.face_normal_x,
                            &g.face_normal_y, &g.face_normal_z, &g.face_area})
                v->assign(v->size(), nan);
            mesh::compute_mesh_geometry(g);
            w.put("geo_cell_cx", g.cell_centroid_x);
            w.put("geo_cell_cy", g.cell_centroid_y);
            w.put("geo_cell_cz", g.cell_centroid_z);
            w.put("geo_cell_vol", g.cell_volume);
            w.put("geo_face_cx", g.face_centroid_x);
            w.put("geo_face_cy", g.face_centroid_y);
            w.put("geo_face_cz", g.face_centroid_z);
            w.put("geo_face_nx", g.face_normal_x);
            w.put("geo_face_ny", g.face_normal_y);
            w.put("geo_face_nz", g.face_normal_z);
            w.put("geo_face_area", g.face_area);
        }
        const auto bcfg = build_bcs(r);
        if (scalar<std::int32_t>(r, "flow_model") == 1) {
            const solver::compressible::physics::PhysicsStack<solver::compressible::physics::ViscousFlow> phys{
                solver::compressible::physics::ViscousFlow{scalar<double>(r, "prandtl")}, {}};
            dispatch_recon(r, mp, bcfg, phys, w);
        } else {
            const solver::compressible::physics::PhysicsStack<solver::compressible::physics::InviscidFlow> phys{
                solver::compressible::physics::InviscidFlow{}, {}};
            dispatch_recon(r, mp, bcfg, phys, w);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_harness: %s\n", e.what());
        rc = 1;
    }
    MPI_Finalize();
    return rc;
}
