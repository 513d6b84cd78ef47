// Host-side unstructured mesh for one rank that owns every cell (no MPI ghosts).
//
// Field-for-field this is the rank-local mesh the reference solver loads from
// its container (include/cfd/mesh/localmesh.hpp:17-102, MeshPart) with
// nprocs == 1: owned cells [0, n_cells), faces in the reference's 3-zone order
// (internal faces sorted by (owner, neigh), then boundary faces sorted by
// (patch, owner); src/mesh/reorder.cpp:422-450), outward normals from owner to
// neighbour, and boundary patches stored as contiguous face ranges.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace cfdg {

enum class CellKind : std::uint8_t { Tet = 0, Pyra = 1, Prism = 2, Hexa = 3, Tri = 4, Quad = 5 };

struct Mesh {
    std::int32_t n_cells = 0;
    std::int32_t n_nodes = 0;
    std::int32_t n_faces = 0;
    std::int32_t n_inner_faces = 0;     // faces with a neighbour cell: [0, n_inner_faces)
    std::int32_t n_internal_faces = 0;  // equal to n_inner_faces for a single rank

    double bbox_lo[3] = {0, 0, 0};
    double bbox_hi[3] = {0, 0, 0};

    // cells
    std::vector<std::uint8_t> cell_type;
    std::vector<std::int64_t> cell_gid;
    std::vector<std::int32_t> cell_nodes_off;  // n_cells + 1
    std::vector<std::int32_t> cell_nodes;
    std::vector<double> cell_cx, cell_cy, cell_cz, cell_vol;

    // nodes
    std::vector<std::int64_t> node_gid;
    std::vector<double> node_x, node_y, node_z;

    // faces
    std::vector<std::int32_t> face_owner, face_neigh, face_patch;
    std::vector<std::uint8_t> face_type;
    std::vector<std::int32_t> face_nodes_off;  // n_faces + 1
    std::vector<std::int32_t> face_nodes;
    std::vector<double> face_cx, face_cy, face_cz;
    std::vector<double> face_nx, face_ny, face_nz, face_area;

    // patches (faces of patch p form one contiguous boundary range)
    std::vector<std::string> patch_names;
    std::vector<std::string> patch_cgns_types;  // CGNS BC type names of the container ("UserDefined" by default)

    std::int32_t n_bfaces() const noexcept { return n_faces - n_inner_faces; }
};

// Parameters of the synthetic structured-to-unstructured generators.
struct MeshSpec {
    int kind = 0;             // 0: hexahedra, 1: six Kuhn tetrahedra per hexahedron
    int ni = 1, nj = 1, nk = 1;
    double x0 = 0, x1 = 1, y0 = 0, y1 = 1, z0 = 0, z1 = 1;
    double y_ratio = 1.0;     // geometric growth of the y spacing away from y0
    double ramp_x = 0.0;      // bottom nodes with x > ramp_x are raised by (x - ramp_x) tan(ramp_deg)
    double ramp_deg = 0.0;    //   and the lift decays linearly to zero at the top (ramp_deg == 0: off)
    double bump_h = 0.0;      // bottom Gaussian bump h exp(-((x - bump_x) / bump_w)^2), decaying to the top
    double bump_w = 1.0;
    double bump_x = 0.0;
    double jitter = 0.0;      // interior nodes move by U(-jitter, jitter) x local spacing per axis
    std::uint64_t seed = 0;
    int split_ymin = 0;       // split the y = y0 side into two patches at x = split_ymin_x
    double split_ymin_x = 0.0;
    int reorder = 0;          // 0: natural i-j-k order, 1: reverse Cuthill-McKee
};

// Builds the mesh: nodes, cells, unique faces with outward owner->neigh
// normals, patch tagging, geometry, optional RCM renumbering and the 3-zone
// face order.  Throws std::runtime_error on invalid input.
Mesh generate_mesh(const MeshSpec& spec);

// Volumes / centroids of cells and areas / normals / centroids of faces,
// computed from the nodes with the tetrahedral / triangular decompositions of
// src/mesh/geometry.cpp:24-150 and 301-498.
void compute_geometry(Mesh& m);

// Reverse Cuthill-McKee order (new -> old) of the face dual graph, following
// src/mesh/reorder.cpp:88-188.
std::vector<std::int32_t> rcm_order(const Mesh& m);

// Renumbers cells by new2old and re-sorts faces into the 3-zone order
// (src/mesh/reorder.cpp:306-576 restricted to one rank).
void apply_cell_permutation(Mesh& m, const std::vector<std::int32_t>& new2old);

// Structural checks (sizes, ranges, zone order, patch contiguity, positive
// volumes, outward normals).  Returns an empty string when the mesh is sane.
std::string check_mesh(const Mesh& m);

}  // namespace cfdg
