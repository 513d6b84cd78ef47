// Host-side, setup-time structures derived from the mesh: cell-centric CSR
// lists, least-squares stencil coefficients, viscous face distances, the BSR
// pattern of the Jacobian and the row groups of the parallel SGS sweeps.
#pragma once

#include <cstdint>
#include <vector>

#include "mesh.hpp"

namespace cfdg {

struct Csr {
    std::vector<std::int32_t> off;  // rows + 1
    std::vector<std::int32_t> idx;
};

// Faces of each cell in ascending face index (aux_connectivity.cpp:11-86).
Csr cell_faces(const Mesh& m);
// Face neighbours of each cell, ascending (aux_connectivity.cpp:88-187).
Csr cell_cells_by_face(const Mesh& m);
// Vertex-sharing neighbours of each cell, ascending (aux_connectivity.cpp:348-445).
Csr cell_cells_by_node(const Mesh& m);

// Weighted least-squares coefficients C_j = A^-1 w_j dr_j, stored as three
// planes of n_edges (gradient_operator.cpp:622-707).
std::vector<double> lsq_coefficients(const Mesh& m, const Csr& stencil);

// Owner->neighbour centroid vectors (mirrored 2 (x_f - x_c) at boundary faces)
// and their inverse lengths (aux_geometry.cpp:11-51, 120-156).
void face_distances(const Mesh& m, std::vector<double>& dx, std::vector<double>& dy, std::vector<double>& dz,
                    std::vector<double>& dinv);

struct BsrPattern {
    std::vector<std::int32_t> row_ptr, cols, diag;
};
// Row i = sort({i} U face neighbours of i) (bsr_helpers.hpp:16-59).
BsrPattern bsr_pattern(const Csr& ccf, std::int32_t n);

// Row groups for the parallel Gauss-Seidel sweeps: rows of one group have no
// matrix coupling, so they can be updated together.
struct RowGroups {
    std::vector<std::int32_t> rows;  // rows ordered by group, ascending inside a group
    std::vector<std::int32_t> off;   // group offsets
};
// Greedy colouring in row order: colour(i) = smallest colour unused by the
// already-coloured (lower-index) neighbours.
std::vector<std::int32_t> greedy_colours(const Csr& ccf, std::int32_t n, std::int32_t& n_colours);
RowGroups groups_from_labels(const std::vector<std::int32_t>& label, std::int32_t n_labels, bool reverse);
// Dependency levels of the natural-order forward sweep (level = 1 + max level
// of lower-index neighbours) and of the backward sweep (higher-index ones).
std::vector<std::int32_t> forward_levels(const Csr& ccf, std::int32_t n, std::int32_t& n_levels);
std::vector<std::int32_t> backward_levels(const Csr& ccf, std::int32_t n, std::int32_t& n_levels);

}  // namespace cfdg
