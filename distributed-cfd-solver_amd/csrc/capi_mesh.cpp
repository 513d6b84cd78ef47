// C ABI of the host-side mesh objects (include/cfdg.h, cfdg_mesh_*).
#include <algorithm>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>

#include <hip/hip_runtime.h>

#include "cfdg.h"
#include "capi_internal.hpp"
#include "kernels_mesh.hpp"
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

namespace {

// Device buffer of one geometry computation, freed on scope exit.
template <class T>
struct DevArray {
    T* p = nullptr;
    explicit DevArray(std::size_t n) {
        if (n > 0 && hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T)) != hipSuccess)
            throw CfdgError(CFDG_ERR_HIP, "hipMalloc failed");
    }
    DevArray(const std::vector<T>& h) : DevArray(h.size()) {
        if (!h.empty() && hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            throw CfdgError(CFDG_ERR_HIP, "hipMemcpy to device failed");
    }
    ~DevArray() {
        if (p) (void)hipFree(p);
    }
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
    void get(std::vector<T>& h) const {
        if (!h.empty() && hipMemcpy(h.data(), p, h.size() * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
            throw CfdgError(CFDG_ERR_HIP, "hipMemcpy to host failed");
    }
};

void geometry_on_device(cfdg::Mesh& m) {
    const auto nc = static_cast<std::size_t>(m.n_cells), nf = static_cast<std::size_t>(m.n_faces);
    for (std::size_t c = 0; c < nc; ++c) {
        const int n = m.cell_nodes_off[c + 1] - m.cell_nodes_off[c];
        const bool tet = m.cell_type[c] == static_cast<std::uint8_t>(cfdg::CellKind::Tet);
        const bool hex = m.cell_type[c] == static_cast<std::uint8_t>(cfdg::CellKind::Hexa);
        if (!((tet && n == 4) || (hex && n == 8)))
            throw CfdgError(CFDG_ERR_MESH, "cell " + std::to_string(c) + ": only tetrahedra and hexahedra");
    }
    for (std::size_t f = 0; f < nf; ++f) {
        const int n = m.face_nodes_off[f + 1] - m.face_nodes_off[f];
        if (n < 3 || n > 8) throw CfdgError(CFDG_ERR_MESH, "face " + std::to_string(f) + " has " + std::to_string(n) + " nodes");
    }
    const auto nn = static_cast<std::int64_t>(m.node_x.size());
    if (m.node_y.size() != m.node_x.size() || m.node_z.size() != m.node_x.size() ||
        m.cell_nodes.size() != static_cast<std::size_t>(m.cell_nodes_off[nc]) ||
        m.face_nodes.size() != static_cast<std::size_t>(m.face_nodes_off[nf]))
        throw CfdgError(CFDG_ERR_MESH, "node lists and offsets disagree");
    for (const auto* list : {&m.cell_nodes, &m.face_nodes})
        for (const int n : *list)
            if (n < 0 || n >= nn) throw CfdgError(CFDG_ERR_MESH, "node index " + std::to_string(n) + " out of range");
    DevArray<double> x(m.node_x), y(m.node_y), z(m.node_z);
    DevArray<int> coff(m.cell_nodes_off), cn(m.cell_nodes), foff(m.face_nodes_off), fn(m.face_nodes);
    DevArray<std::uint8_t> ct(m.cell_type);
    DevArray<double> ccx(nc), ccy(nc), ccz(nc), cvol(nc), fcx(nf), fcy(nf), fcz(nf), fnx(nf), fny(nf), fnz(nf),
        farea(nf);
    const cfdg::GeomArgs a{m.n_cells, m.n_faces, x.p, y.p, z.p, coff.p, cn.p, ct.p, foff.p, fn.p,
                           ccx.p, ccy.p, ccz.p, cvol.p, fcx.p, fcy.p, fcz.p, fnx.p, fny.p, fnz.p, farea.p};
    cfdg::launch_mesh_geometry(a, nullptr);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        throw CfdgError(CFDG_ERR_HIP, "geometry kernels failed");
    std::vector<double> vol(nc), area(nf);
    cvol.get(vol);
    farea.get(area);
    for (std::size_t c = 0; c < nc; ++c)
        if (!(vol[c] > 1e-15))
            throw CfdgError(CFDG_ERR_MESH, "degenerate cell " + std::to_string(c) + " volume " + std::to_string(vol[c]));
    for (std::size_t f = 0; f < nf; ++f)
        if (!(area[f] > 1e-15))
            throw CfdgError(CFDG_ERR_MESH, "degenerate face " + std::to_string(f) + " area " + std::to_string(area[f]));
    m.cell_vol = std::move(vol);
    m.face_area = std::move(area);
    for (auto* v : {&m.cell_cx, &m.cell_cy, &m.cell_cz}) v->resize(nc);
    for (auto* v : {&m.face_cx, &m.face_cy, &m.face_cz, &m.face_nx, &m.face_ny, &m.face_nz}) v->resize(nf);
    ccx.get(m.cell_cx);
    ccy.get(m.cell_cy);
    ccz.get(m.cell_cz);
    fcx.get(m.face_cx);
    fcy.get(m.face_cy);
    fcz.get(m.face_cz);
    fnx.get(m.face_nx);
    fny.get(m.face_ny);
    fnz.get(m.face_nz);
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
        mesh->conn.reset();
        return CFDG_OK;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "cfdg_mesh_reorder_rcm: %s\n", e.what());
        return CFDG_ERR_MESH;
    }
}


int cfdg_mesh_geometry_device(cfdg_mesh* mesh) {
    if (!mesh) return CFDG_ERR_ARG;
    try {
        cfdg::Mesh m = mesh->m;  // the mesh is left unchanged on error
        geometry_on_device(m);
        mesh->m = std::move(m);
        return CFDG_OK;
    } catch (const CfdgError& e) {
        std::fprintf(stderr, "cfdg_mesh_geometry_device: %s\n", e.what());
        return e.code;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "cfdg_mesh_geometry_device: %s\n", e.what());
        return CFDG_ERR_MESH;
    }
}

int cfdg_mesh_reorder_rcm_device(cfdg_mesh* mesh) {
    if (!mesh) return CFDG_ERR_ARG;
    try {
        cfdg::Mesh m = mesh->m;  // the mesh is left unchanged on error
        cfdg::apply_cell_permutation(m, cfdg::rcm_order_device(m));
        mesh->m = std::move(m);
        mesh->conn.reset();
        return CFDG_OK;
    } catch (const cfdg::DeviceError& e) {
        std::fprintf(stderr, "cfdg_mesh_reorder_rcm_device: %s\n", e.what());
        return CFDG_ERR_HIP;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "cfdg_mesh_reorder_rcm_device: %s\n", e.what());
        return CFDG_ERR_MESH;
    }
}

int cfdg_mesh_connectivity(cfdg_mesh* mesh, int32_t on_device, int32_t node_stencil) {
    if (!mesh) return CFDG_ERR_ARG;
    try {
        const bool ns = node_stencil != 0;
        mesh->conn = std::make_shared<const cfdg::MeshConnectivity>(
            on_device ? cfdg::device_connectivity(mesh->m, ns) : cfdg::host_connectivity(mesh->m, ns));
        return CFDG_OK;
    } catch (const cfdg::DeviceError& e) {
        std::fprintf(stderr, "cfdg_mesh_connectivity: %s\n", e.what());
        return CFDG_ERR_HIP;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "cfdg_mesh_connectivity: %s\n", e.what());
        return CFDG_ERR_MESH;
    }
}

int cfdg_mesh_get_connectivity(const cfdg_mesh* mesh, cfdg_mesh_connectivity_view* v) {
    if (!mesh || !v) return CFDG_ERR_ARG;
    if (!mesh->conn) return CFDG_ERR_STATE;
    const cfdg::MeshConnectivity& c = *mesh->conn;
    v->n_cells = mesh->m.n_cells;
    v->has_node_stencil = c.has_node_stencil ? 1 : 0;
    v->cell_faces_off = c.cell_faces.off.data();
    v->cell_faces = c.cell_faces.idx.data();
    v->cell_cells_face_off = c.cell_cells_face.off.data();
    v->cell_cells_face = c.cell_cells_face.idx.data();
    v->cell_cells_node_off = c.has_node_stencil ? c.cell_cells_node.off.data() : nullptr;
    v->cell_cells_node = c.has_node_stencil ? c.cell_cells_node.idx.data() : nullptr;
    v->bsr_row_ptr = c.bsr.row_ptr.data();
    v->bsr_cols = c.bsr.cols.data();
    v->bsr_diag = c.bsr.diag.data();
    return CFDG_OK;
}

void cfdg_mesh_free(cfdg_mesh* mesh) { delete mesh; }

}  // extern "C"
