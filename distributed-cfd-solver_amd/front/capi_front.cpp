// C ABI of the host front end (include/cfdg_front.h): configuration files,
// mesh container, and the conversion of a parsed configuration to the
// device's numerics.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "cfdg.h"
#include "cfdg_front.h"
#include "config.hpp"
#include "container.hpp"

namespace {

using cfdg::front::BcDescriptor;
using cfdg::front::SolverConfig;

void put_msg(char* buf, int32_t len, const std::string& s) {
    if (buf && len > 0) std::snprintf(buf, static_cast<std::size_t>(len), "%s", s.c_str());
}

int to_c(const SolverConfig& s, cfdg_solver_config* o, char* err, int32_t errlen) {
    if (s.output_dir.size() >= sizeof(o->output_dir)) {
        put_msg(err, errlen, "config: output directory longer than " + std::to_string(sizeof(o->output_dir) - 1) +
                                 " characters");
        return CFDG_ERR_ARG;
    }
    *o = cfdg_solver_config{};
    o->eos = static_cast<int32_t>(s.eos);
    o->gamma = s.gamma;
    o->gas_constant = s.gas_constant;
    o->prandtl = s.prandtl;
    o->flow_model = static_cast<int32_t>(s.flow_model);
    o->init_rho = s.init_rho;
    o->init_p = s.init_p;
    for (int k = 0; k < 3; ++k) o->init_velocity[k] = s.init_velocity[static_cast<std::size_t>(k)];
    o->gradient = static_cast<int32_t>(s.gradient);
    o->flux = static_cast<int32_t>(s.flux);
    o->reconstruction = static_cast<int32_t>(s.reconstruction);
    o->limiter = static_cast<int32_t>(s.limiter);
    o->venkat_k = s.venkat_k;
    o->turbulence = s.turbulence ? 1 : 0;
    o->nu_inf_ratio = s.nu_inf_ratio;
    o->max_distance_sweeps = s.max_distance_sweeps;
    o->distance_tolerance = s.distance_tolerance;
    o->scheme = static_cast<int32_t>(s.scheme);
    o->cfl = s.cfl;
    o->max_iterations = s.max_iterations;
    o->residual_tolerance = s.residual_tolerance;
    o->time_mode = static_cast<int32_t>(s.time_mode);
    o->dt = s.dt;
    o->max_time_steps = s.max_time_steps;
    o->bdf_order = s.bdf_order;
    o->solver = static_cast<int32_t>(s.solver);
    o->preconditioner = static_cast<int32_t>(s.preconditioner);
    o->rel_tol = s.rel_tol;
    o->abs_tol = s.abs_tol;
    o->max_iter = s.max_iter;
    o->verbosity = static_cast<int32_t>(s.verbosity);
    std::memcpy(o->output_dir, s.output_dir.c_str(), s.output_dir.size() + 1);
    o->field_interval = s.field_interval;
    o->residual_interval = s.residual_interval;
    return CFDG_OK;
}

void bc_to_c(const BcDescriptor& d, cfdg_bc& b) {
    b = cfdg_bc{};
    b.type = static_cast<int32_t>(d.type);
    b.inflow_mode = static_cast<int32_t>(d.inflow_mode);
    b.p = d.p;
    b.t = d.t;
    b.tmp_grad = d.tmp_grad;
    for (int k = 0; k < 3; ++k) {
        b.velocity[k] = d.velocity[static_cast<std::size_t>(k)];
        b.direction[k] = d.direction[static_cast<std::size_t>(k)];
    }
    b.mach = d.mach;
    b.alpha_deg = d.alpha_deg;
    b.beta_deg = d.beta_deg;
}

int read_boundary(const std::string* text, const char* path, const std::vector<std::string>& names, cfdg_bc* bcs,
                  int32_t n_bcs, char* warnings, int32_t wlen, char* err, int32_t errlen) {
    if (!path || !bcs || n_bcs != static_cast<int32_t>(names.size())) {
        put_msg(err, errlen, "config: boundary conditions: n_bcs must equal the patch count");
        return CFDG_ERR_ARG;
    }
    try {
        const cfdg::front::BoundaryConfig bc = text ? cfdg::front::parse_boundary_config_text(*text, path, names)
                                                    : cfdg::front::parse_boundary_config(path, names);
        for (std::size_t p = 0; p < names.size(); ++p) bc_to_c(bc.patches[p], bcs[p]);
        std::string w;
        for (const auto& s : bc.warnings) w += s + "\n";
        put_msg(warnings, wlen, w);
        return CFDG_OK;
    } catch (const cfdg::front::ConfigError& e) {
        put_msg(err, errlen, e.what());
        return CFDG_ERR_CONFIG;
    }
}

// cfdg::Mesh <-> container, one partition
cfdg::front::ContainerMesh to_container(const cfdg::Mesh& m) {
    cfdg::front::ContainerMesh c;
    c.n_cells_g = static_cast<std::uint64_t>(m.n_cells);
    c.n_faces_g = static_cast<std::uint64_t>(m.n_faces);
    c.n_bfaces_g = static_cast<std::uint64_t>(m.n_faces - m.n_inner_faces);
    c.n_nodes_g = static_cast<std::uint64_t>(m.n_nodes);
    for (int k = 0; k < 3; ++k) {  // domain bounding box = that of the nodes
        const std::vector<double>& x = k == 0 ? m.node_x : k == 1 ? m.node_y : m.node_z;
        c.bbox_lo[static_cast<std::size_t>(k)] = x.empty() ? 0.0 : *std::min_element(x.begin(), x.end());
        c.bbox_hi[static_cast<std::size_t>(k)] = x.empty() ? 0.0 : *std::max_element(x.begin(), x.end());
    }
    c.cell_type = m.cell_type;
    c.cell_gid = m.cell_gid;
    c.cell_donor.assign(static_cast<std::size_t>(m.n_cells), -1);
    c.cell_volume = m.cell_vol;
    c.cell_cx = m.cell_cx;
    c.cell_cy = m.cell_cy;
    c.cell_cz = m.cell_cz;
    c.cell_nodes_off = m.cell_nodes_off;
    c.cell_nodes = m.cell_nodes;
    c.face_owner = m.face_owner;
    c.face_neigh = m.face_neigh;
    c.face_patch = m.face_patch;
    c.face_type = m.face_type;
    c.face_area = m.face_area;
    c.face_nx = m.face_nx;
    c.face_ny = m.face_ny;
    c.face_nz = m.face_nz;
    c.face_cx = m.face_cx;
    c.face_cy = m.face_cy;
    c.face_cz = m.face_cz;
    c.face_nodes_off = m.face_nodes_off;
    c.face_nodes = m.face_nodes;
    c.node_gid = m.node_gid;
    c.node_x = m.node_x;
    c.node_y = m.node_y;
    c.node_z = m.node_z;
    c.send_offsets = {0};
    c.recv_offsets = {0};
    c.patch_names = m.patch_names;
    c.patch_cgns_types = m.patch_cgns_types;
    c.patch_cgns_types.resize(m.patch_names.size(), "UserDefined");
    // boundary faces of patch p: a contiguous range in the 3-zone order
    const std::size_t np = m.patch_names.size();
    c.patch_face_offsets.assign(np + 1, 0);
    for (std::int32_t f = m.n_inner_faces; f < m.n_faces; ++f)
        ++c.patch_face_offsets[static_cast<std::size_t>(m.face_patch[static_cast<std::size_t>(f)]) + 1];
    for (std::size_t p = 0; p < np; ++p) c.patch_face_offsets[p + 1] += c.patch_face_offsets[p];
    std::vector<std::int32_t> fill(c.patch_face_offsets.begin(), c.patch_face_offsets.end() - 1);
    c.patch_faces.assign(static_cast<std::size_t>(m.n_faces - m.n_inner_faces), 0);
    for (std::int32_t f = m.n_inner_faces; f < m.n_faces; ++f)
        c.patch_faces[static_cast<std::size_t>(fill[static_cast<std::size_t>(m.face_patch[static_cast<std::size_t>(f)])]++)] = f;
    return c;
}

}  // namespace

extern "C" {

int cfdg_config_read_solver(const char* path, cfdg_solver_config* out, char* err, int32_t errlen) {
    if (!path || !out) return CFDG_ERR_ARG;
    try {
        return to_c(cfdg::front::parse_solver_config(path), out, err, errlen);
    } catch (const cfdg::front::ConfigError& e) {
        put_msg(err, errlen, e.what());
        return CFDG_ERR_CONFIG;
    }
}

int cfdg_config_read_solver_text(const char* text, const char* path, cfdg_solver_config* out, char* err,
                                 int32_t errlen) {
    if (!text || !path || !out) return CFDG_ERR_ARG;
    try {
        return to_c(cfdg::front::parse_solver_config_text(text, path), out, err, errlen);
    } catch (const cfdg::front::ConfigError& e) {
        put_msg(err, errlen, e.what());
        return CFDG_ERR_CONFIG;
    }
}

int cfdg_config_read_boundary(const char* path, const cfdg_mesh* mesh, cfdg_bc* bcs, int32_t n_bcs, char* warnings,
                              int32_t wlen, char* err, int32_t errlen) {
    if (!mesh) return CFDG_ERR_ARG;
    return read_boundary(nullptr, path, mesh->m.patch_names, bcs, n_bcs, warnings, wlen, err, errlen);
}

int cfdg_config_read_boundary_text(const char* text, const char* path, const char* const* patch_names,
                                   int32_t n_patches, cfdg_bc* bcs, char* warnings, int32_t wlen, char* err,
                                   int32_t errlen) {
    if (!text || n_patches < 0 || (n_patches > 0 && !patch_names)) return CFDG_ERR_ARG;
    std::vector<std::string> names;
    for (int32_t p = 0; p < n_patches; ++p) names.emplace_back(patch_names[p] ? patch_names[p] : "");
    const std::string t(text);
    return read_boundary(&t, path, names, bcs, n_patches, warnings, wlen, err, errlen);
}

int cfdg_config_to_numerics(const cfdg_solver_config* cfg, cfdg_numerics* num, cfdg_linparams* lp, char* err,
                            int32_t errlen) {
    if (!cfg || !num || !lp) return CFDG_ERR_ARG;
    const auto reject = [&](const std::string& what) {
        put_msg(err, errlen, what);
        return CFDG_ERR_CONFIG;
    };
    cfdg_numerics n{};
    if (cfg->eos != CFDG_CFG_EOS_IDEAL_GAS) return reject("dispatch: unknown EOS model");
    if (cfg->flux != CFDG_CFG_FLUX_HLLC) return reject("dispatch: unknown flux scheme");
    if (cfg->flow_model == CFDG_CFG_FLOW_VISCOUS)
        n.flow_model = CFDG_VISCOUS_FLOW;
    else if (cfg->flow_model == CFDG_CFG_FLOW_INVISCID)
        n.flow_model = CFDG_INVISCID_FLOW;
    else
        return reject("dispatch: unknown Flow Model model");
    n.gamma = cfg->gamma;
    n.gas_constant = cfg->gas_constant;
    n.prandtl = cfg->prandtl;
    n.prandtl_turb = 0.9;  // kTurbPrandtl (include/cfd/core/types.hpp:44)
    if (cfg->reconstruction == CFDG_CFG_RECON_FIRST_ORDER) {
        n.reconstruction = CFDG_FIRST_ORDER;
        n.limiter = CFDG_LIMITER_VENKAT;  // not used by first order
    } else if (cfg->reconstruction == CFDG_CFG_RECON_MUSCL) {
        n.reconstruction = CFDG_MUSCL;
        switch (cfg->limiter) {
            case CFDG_CFG_LIMITER_VENKATAKRISHNAN: n.limiter = CFDG_LIMITER_VENKAT; break;
            case CFDG_CFG_LIMITER_BARTH_JESPERSEN: n.limiter = CFDG_LIMITER_BARTH; break;
            case CFDG_CFG_LIMITER_VAN_ALBADA: n.limiter = CFDG_LIMITER_VAN_ALBADA; break;
            default: return reject("dispatch: invalid multidimensional limiter type");
        }
    } else if (cfg->reconstruction == CFDG_CFG_RECON_MUSCL_DIRECTIONAL) {
        return reject("dispatch: MUSCL_DIRECTIONAL reconstruction is not available on the device path");
    } else {
        return reject("dispatch: unknown reconstruction scheme");
    }
    n.venkat_k = cfg->venkat_k;
    switch (cfg->gradient) {
        case CFDG_CFG_GRADIENT_GREEN_GAUSS_FACE: n.gradient = CFDG_GREEN_GAUSS_FACE; break;
        case CFDG_CFG_GRADIENT_GREEN_GAUSS_CELL: n.gradient = CFDG_GREEN_GAUSS_CELL; break;
        case CFDG_CFG_GRADIENT_LEAST_SQUARES_CELL_NODE: n.gradient = CFDG_LEAST_SQUARES_NODE; break;
        default: return reject("dispatch: the face-stencil least-squares gradient is not available on the device path");
    }
    n.cfl = cfg->cfl;
    cfdg_linparams l{};
    l.preconditioner = cfg->preconditioner == CFDG_CFG_PRECOND_SGS ? CFDG_PRECOND_SGS : CFDG_PRECOND_NONE;
    l.sgs_order = CFDG_SGS_MULTICOLOUR;
    l.rel_tol = cfg->rel_tol;
    l.abs_tol = cfg->abs_tol;
    l.max_iter = cfg->max_iter;
    *num = n;
    *lp = l;
    return CFDG_OK;
}

int cfdg_container_read(const char* path, cfdg_mesh** out, char* err, int32_t errlen) {
    if (!path || !out) return CFDG_ERR_ARG;
    *out = nullptr;
    try {
        const cfdg::front::ContainerMesh c = cfdg::front::read_container(path);
        if (c.n_own != c.n_cells()) {
            put_msg(err, errlen, "container: " + std::to_string(c.n_cells() - c.n_own) +
                                     " halo cells in a one-rank partition");
            return CFDG_ERR_IO;
        }
        auto* mesh = new cfdg_mesh();
        cfdg::Mesh& m = mesh->m;
        m.n_cells = c.n_cells();
        m.n_nodes = c.n_nodes();
        m.n_faces = c.n_faces();
        m.n_inner_faces = c.n_inner_faces;
        m.n_internal_faces = c.n_internal_faces;
        for (int k = 0; k < 3; ++k) {
            m.bbox_lo[k] = c.bbox_lo[static_cast<std::size_t>(k)];
            m.bbox_hi[k] = c.bbox_hi[static_cast<std::size_t>(k)];
        }
        m.cell_type = c.cell_type;
        m.cell_gid = c.cell_gid;
        m.cell_nodes_off = c.cell_nodes_off;
        m.cell_nodes = c.cell_nodes;
        m.cell_cx = c.cell_cx;
        m.cell_cy = c.cell_cy;
        m.cell_cz = c.cell_cz;
        m.cell_vol = c.cell_volume;
        m.node_gid = c.node_gid;
        m.node_x = c.node_x;
        m.node_y = c.node_y;
        m.node_z = c.node_z;
        m.face_owner = c.face_owner;
        m.face_neigh = c.face_neigh;
        m.face_patch = c.face_patch;
        m.face_type = c.face_type;
        m.face_nodes_off = c.face_nodes_off;
        m.face_nodes = c.face_nodes;
        m.face_cx = c.face_cx;
        m.face_cy = c.face_cy;
        m.face_cz = c.face_cz;
        m.face_nx = c.face_nx;
        m.face_ny = c.face_ny;
        m.face_nz = c.face_nz;
        m.face_area = c.face_area;
        m.patch_names = c.patch_names;
        m.patch_cgns_types = c.patch_cgns_types;
        for (const auto n : m.cell_nodes)
            if (n < 0 || n >= m.n_nodes) {
                delete mesh;
                put_msg(err, errlen, "container: cell node index out of range");
                return CFDG_ERR_MESH;
            }
        const std::string e = cfdg::check_mesh(m);
        if (!e.empty()) {
            delete mesh;
            put_msg(err, errlen, "container: " + e);
            return CFDG_ERR_MESH;
        }
        *out = mesh;
        return CFDG_OK;
    } catch (const cfdg::front::ContainerError& e) {
        put_msg(err, errlen, e.what());
        return CFDG_ERR_IO;
    } catch (const std::exception& e) {
        put_msg(err, errlen, e.what());
        return CFDG_ERR_IO;
    }
}

int cfdg_container_write(const char* path, const cfdg_mesh* mesh, char* err, int32_t errlen) {
    if (!path || !mesh) return CFDG_ERR_ARG;
    try {
        cfdg::front::write_container(path, to_container(mesh->m));
        return CFDG_OK;
    } catch (const std::exception& e) {
        put_msg(err, errlen, e.what());
        return CFDG_ERR_IO;
    }
}

}  // extern "C"
