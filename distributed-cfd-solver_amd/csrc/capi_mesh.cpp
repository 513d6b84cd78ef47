// C ABI of the host-side mesh objects (include/cfdg.h, cfdg_mesh_*).
#include <algorithm>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "cfdg.h"
#include "capi_internal.hpp"
#include "mesh.hpp"

namespace {

using cfdg::CfdgError;

cfdg::Mesh mesh_from_view(const cfdg_mesh_view& v) {
    if (v.n_cells <= 0 || v.n_faces <= 0 || v.n_inner_faces < 0 || v.n_inner_faces > v.n_faces || v.n_patches < 0)
        throw CfdgError(CFDG_ERR_ARG, "mesh view: invalid sizes");
    cfdg::Mesh m;
    const auto nc = static_cast<std::size_t>(v.n_cells), nf = static_cast<std::size_t>(v.n_faces),
               nn = static_cast<std::size_t>(v.n_nodes);
    m.n_cells = v.n_cells;
    m.n_nodes = v.n_nodes;
    m.n_faces = v.n_faces;
    m.n_inner_faces = v.n_inner_faces;
    m.n_internal_faces = v.n_inner_faces;
    m.cell_type.assign(v.cell_type, v.cell_type + nc);
    m.cell_gid.resize(nc);
    for (std::size_t i = 0; i < nc; ++i) m.cell_gid[i] = static_cast<std::int64_t>(i);
    m.cell_nodes_off.assign(v.cell_nodes_off, v.cell_nodes_off + nc + 1);
    m.cell_nodes.assign(v.cell_nodes, v.cell_nodes + m.cell_nodes_off.back());
    m.cell_cx.assign(v.cell_cx, v.cell_cx + nc);
    m.cell_cy.assign(v.cell_cy, v.cell_cy + nc);
    m.cell_cz.assign(v.cell_cz, v.cell_cz + nc);
    m.cell_vol.assign(v.cell_vol, v.cell_vol + nc);
    m.node_x.assign(v.node_x, v.node_x + nn);
    m.node_y.assign(v.node_y, v.node_y + nn);
    m.node_z.assign(v.node_z, v.node_z + nn);
    m.node_gid.resize(nn);
    for (std::size_t i = 0; i < nn; ++i) m.node_gid[i] = static_cast<std::int64_t>(i);
    m.face_owner.assign(v.face_owner, v.face_owner + nf);
    m.face_neigh.assign(v.face_neigh, v.face_neigh + nf);
    m.face_patch.assign(v.face_patch, v.face_patch + nf);
    m.face_type.assign(v.face_type, v.face_type + nf);
    m.face_nodes_off.assign(v.face_nodes_off, v.face_nodes_off + nf + 1);
    m.face_nodes.assign(v.face_nodes, v.face_nodes + m.face_nodes_off.back());
    m.face_cx.assign(v.face_cx, v.face_cx + nf);
    m.face_cy.assign(v.face_cy, v.face_cy + nf);
    m.face_cz.assign(v.face_cz, v.face_cz + nf);
    m.face_nx.assign(v.face_nx, v.face_nx + nf);
    m.face_ny.assign(v.face_ny, v.face_ny + nf);
    m.face_nz.assign(v.face_nz, v.face_nz + nf);
    m.face_area.assign(v.face_area, v.face_area + nf);
    for (std::int32_t p = 0; p < v.n_patches; ++p) m.patch_names.push_back("patch" + std::to_string(p));
    m.patch_cgns_types.assign(static_cast<std::size_t>(v.n_patches), "UserDefined");
    for (std::size_t i = 0; i < nc; ++i) {
        if (m.cell_nodes_off[i + 1] < m.cell_nodes_off[i]) throw CfdgError(CFDG_ERR_MESH, "cell node offsets");
    }
    for (const auto n : m.cell_nodes)
        if (n < 0 || n >= v.n_nodes) throw CfdgError(CFDG_ERR_MESH, "cell node index out of range");
    const std::string e = cfdg::check_mesh(m);
    if (!e.empty()) throw CfdgError(CFDG_ERR_MESH, e);
    return m;
}

}  // namespace

extern "C" {

int cfdg_mesh_generate(const cfdg_mesh_spec* spec, cfdg_mesh** out) {
    if (!spec || !out) return CFDG_ERR_ARG;
    *out = nullptr;
    try {
        cfdg::MeshSpec s;
        s.kind = spec->kind;
        s.ni = spec->ni;
        s.nj = spec->nj;
        s.nk = spec->nk;
        s.x0 = spec->x0;
        s.x1 = spec->x1;
        s.y0 = spec->y0;
        s.y1 = spec->y1;
        s.z0 = spec->z0;
        s.z1 = spec->z1;
        s.y_ratio = spec->y_ratio;
        s.ramp_x = spec->ramp_x;
        s.ramp_deg = spec->ramp_deg;
        s.bump_h = spec->bump_h;
        s.bump_w = spec->bump_w;
        s.bump_x = spec->bump_x;
        s.jitter = spec->jitter;
        s.seed = spec->seed;
        s.split_ymin = spec->split_ymin;
        s.split_ymin_x = spec->split_ymin_x;
        s.reorder = spec->reorder;
        auto* m = new cfdg_mesh();
        try {
            m->m = cfdg::generate_mesh(s);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
        return CFDG_OK;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "cfdg_mesh_generate: %s\n", e.what());
        return CFDG_ERR_MESH;
    }
}

int cfdg_mesh_from_view(const cfdg_mesh_view* view, cfdg_mesh** out) {
    if (!view || !out) return CFDG_ERR_ARG;
    *out = nullptr;
    try {
        auto* m = new cfdg_mesh();
        try {
            m->m = mesh_from_view(*view);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
        return CFDG_OK;
    } catch (const CfdgError& e) {
        std::fprintf(stderr, "cfdg_mesh_from_view: %s\n", e.what());
        return e.code;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "cfdg_mesh_from_view: %s\n", e.what());
        return CFDG_ERR_MESH;
    }
}

int cfdg_mesh_get_view(const cfdg_mesh* mesh, cfdg_mesh_view* v) {
    if (!mesh || !v) return CFDG_ERR_ARG;
    const cfdg::Mesh& m = mesh->m;
    v->n_cells = m.n_cells;
    v->n_nodes = m.n_nodes;
    v->n_faces = m.n_faces;
    v->n_inner_faces = m.n_inner_faces;
    v->n_patches = static_cast<std::int32_t>(m.patch_names.size());
    v->cell_type = m.cell_type.data();
    v->cell_nodes_off = m.cell_nodes_off.data();
    v->cell_nodes = m.cell_nodes.data();
    v->cell_cx = m.cell_cx.data();
    v->cell_cy = m.cell_cy.data();
    v->cell_cz = m.cell_cz.data();
    v->cell_vol = m.cell_vol.data();
    v->node_x = m.node_x.data();
    v->node_y = m.node_y.data();
    v->node_z = m.node_z.data();
    v->face_owner = m.face_owner.data();
    v->face_neigh = m.face_neigh.data();
    v->face_patch = m.face_patch.data();
    v->face_type = m.face_type.data();
    v->face_nodes_off = m.face_nodes_off.data();
    v->face_nodes = m.face_nodes.data();
    v->face_cx = m.face_cx.data();
    v->face_cy = m.face_cy.data();
    v->face_cz = m.face_cz.data();
    v->face_nx = m.face_nx.data();
    v->face_ny = m.face_ny.data();
    v->face_nz = m.face_nz.data();
    v->face_area = m.face_area.data();
    return CFDG_OK;
}

int cfdg_mesh_patch_name(const cfdg_mesh* mesh, int32_t patch, char* buf, int32_t buflen) {
    if (!mesh || !buf || buflen <= 0) return CFDG_ERR_ARG;
    if (patch < 0 || patch >= static_cast<int32_t>(mesh->m.patch_names.size())) return CFDG_ERR_ARG;
    std::snprintf(buf, static_cast<std::size_t>(buflen), "%s", mesh->m.patch_names[static_cast<std::size_t>(patch)].c_str());
    return CFDG_OK;
}

int cfdg_mesh_patch_cgns_type(const cfdg_mesh* mesh, int32_t patch, char* buf, int32_t buflen) {
    if (!mesh || !buf || buflen <= 0) return CFDG_ERR_ARG;
    if (patch < 0 || patch >= static_cast<int32_t>(mesh->m.patch_cgns_types.size())) return CFDG_ERR_ARG;
    std::snprintf(buf, static_cast<std::size_t>(buflen), "%s",
                  mesh->m.patch_cgns_types[static_cast<std::size_t>(patch)].c_str());
    return CFDG_OK;
}

int cfdg_mesh_set_patch(cfdg_mesh* mesh, int32_t patch, const char* name, const char* cgns_type) {
    if (!mesh) return CFDG_ERR_ARG;
    if (patch < 0 || patch >= static_cast<int32_t>(mesh->m.patch_names.size())) return CFDG_ERR_ARG;
    if (name) mesh->m.patch_names[static_cast<std::size_t>(patch)] = name;
    if (cgns_type) mesh->m.patch_cgns_types[static_cast<std::size_t>(patch)] = cgns_type;
    return CFDG_OK;
}

int cfdg_mesh_cell_gid(const cfdg_mesh* mesh, int64_t* gid) {
    if (!mesh || !gid) return CFDG_ERR_ARG;
    std::copy(mesh->m.cell_gid.begin(), mesh->m.cell_gid.end(), gid);
    return CFDG_OK;
}

int cfdg_mesh_reorder_rcm(cfdg_mesh* mesh) {
    if (!mesh) return CFDG_ERR_ARG;
    try {
        cfdg::apply_cell_permutation(mesh->m, cfdg::rcm_order(mesh->m));
        return CFDG_OK;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "cfdg_mesh_reorder_rcm: %s\n", e.what());
        return CFDG_ERR_MESH;
    }
}

void cfdg_mesh_free(cfdg_mesh* mesh) { delete mesh; }

}  // extern "C"
