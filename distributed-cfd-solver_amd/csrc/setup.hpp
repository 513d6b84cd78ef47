// Host-side, setup-time structures derived from the mesh: cell-centric CSR
// lists, least-squares stencil coefficients, viscous face distances, the BSR
// pattern of the Jacobian and the row groups of the parallel SGS sweeps.
#pragma once

#include <cstdint>
#include <stdexcept>
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

// Distance of every cell centroid to the nearest wall face centroid along the
// stencil graph (aux_geometry.cpp:545-697); wall_patch[p] marks the wall patches.
std::vector<double> wall_distance(const Mesh& m, const Csr& stencil, const std::vector<char>& wall_patch,
                                  int max_sweeps, double rel_tol);

struct BsrPattern {
    std::vector<std::int32_t> row_ptr, cols, diag;
};
// Row i = sort({i} U face neighbours of i) (bsr_helpers.hpp:16-59).
BsrPattern bsr_pattern(const Csr& ccf, std::int32_t n);

// The cell connectivity a context is built from: the faces of each cell,
// its face and vertex neighbours and the BSR pattern of the Jacobian
// (MeshAuxConnectivity, aux_connectivity.cpp:11-187, 348-445, and
// bsr_helpers.hpp:16-59).  The vertex stencil is built only on request
// (node-LSQ gradients and the wall distance use it).
struct MeshConnectivity {
    Csr cell_faces, cell_cells_face, cell_cells_node;
    BsrPattern bsr;
    bool has_node_stencil = false;
};
// The builders above, on the host.
MeshConnectivity host_connectivity(const Mesh& m, bool node_stencil);
// A HIP runtime failure of a device-side set-up step.
struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
// The same arrays built on the current HIP device (kernels_conn.hip), equal
// to host_connectivity's entry for entry; throws DeviceError on a HIP failure
// and std::runtime_error on a mesh whose faces couple a pair of cells twice.
MeshConnectivity device_connectivity(const Mesh& m, bool node_stencil);
// The reverse Cuthill-McKee order of rcm_order (mesh.cpp) computed on the
// current HIP device (kernels_conn.hip): the same new -> old permutation,
// entry for entry; throws DeviceError on a HIP failure.
std::vector<std::int32_t> rcm_order_device(const Mesh& m);

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

// Device cell order: the cells are grouped into compact tiles by recursive
// bisection of the centroids.  Each split cuts the set at the median of one
// coordinate axis, choosing the axis whose median plane crosses the fewest
// faces (on stretched boundary-layer cells this splits along the direction
// with the most cells, not the longest one).  A set becomes a tile once it
// holds at most max_cells cells and at most max_faces faces (interior +
// boundary, counted once).  Tiles come out in depth-first order of the
// bisection tree, so consecutive tiles, and any contiguous range of tiles,
// are spatially compact.  Inside a tile cells keep ascending index order.
struct Tiling {
    std::vector<std::int32_t> new2old;   // device cell -> mesh cell
    std::vector<std::int32_t> old2new;
    std::vector<std::int32_t> tile_off;  // tiles + 1, offsets into the device order
};
Tiling tile_cells(const Mesh& m, const Csr& ccf, const Csr& cf, int max_cells, int max_faces);

// Per-tile face lists for the tiled flux kernel, in device numbering:
//   faces[tile_face_off[t] .. tile_face_off[t+1]) = the faces touching tile t,
//   grouped by which of the face's cells lie in the tile -- both, the owner
//   only (boundary faces included: their ghost is outside every tile), the
//   neighbour only -- and ascending face index within each group, so that the
//   lanes of a wave mostly take the same path; cell_slot[e] (parallel to the
//   device cell->faces list) = position of that face in its tile's list,
//   negated and minus one when the cell is the face's neighbour (the flux
//   enters with a minus sign).  face_other_dev = neighbour device id, or the
//   ghost id (>= cells) of a boundary face.
struct TileFaces {
    std::vector<std::int32_t> tile_face_off;
    std::vector<std::int32_t> faces;
    std::vector<std::int32_t> cell_slot;
};
TileFaces tile_faces(const Tiling& t, const Csr& dev_cf, const std::vector<std::int32_t>& face_owner_dev,
                     const std::vector<std::int32_t>& face_other_dev);

// LDS-staged neighbour stencils for the gradient kernel, in device numbering.
// A tile's LDS slots are its own cells (slot = cell - tile start) followed by
// halo[halo_off[t] .. halo_off[t+1]): the cells outside the tile that its
// stencils reach, ascending.  ccn_slot / ccf_slot give, for every entry of the
// device vertex-neighbour / face-neighbour lists, the neighbour's slot in the
// LDS of the tile that owns the row.  max_slots is the largest tile + halo.
// The slots are stored per tile in chunks of four: entry k of the list of the
// tile's i-th cell at ccn_base[t] + ((k / 4) nt + i) 4 + k % 4 (nt = cells of
// the tile; ccn_base[t+1] - ccn_base[t] = nt times the tile's longest list
// rounded up to a multiple of 4; entries past a cell's own count are padding),
// so that one 8-byte load brings four entries of a cell and the lanes of a
// wave (the tile's consecutive cells) read consecutive chunks.  Likewise
// ccf_slot / ccf_base.
struct TileStencil {
    std::vector<std::int32_t> halo_off;
    std::vector<std::int32_t> halo;
    std::vector<std::uint16_t> ccn_slot;
    std::vector<std::uint16_t> ccf_slot;
    std::vector<std::int32_t> ccn_base;
    std::vector<std::int32_t> ccf_base;
    std::int32_t max_slots = 0;
};
TileStencil tile_stencil(const std::vector<std::int32_t>& tile_off, const Csr& dev_ccn, const Csr& dev_ccf);
// The largest slot count (cells + stencil halo) of any tile, without the slot lists.
std::int32_t tile_max_slots(const std::vector<std::int32_t>& tile_off, const Csr& dev_ccn, const Csr& dev_ccf);

// Splits each tile (a range of device cells) into consecutive sub-ranges whose
// cells plus stencil halo take at most max_slots LDS slots, for meshes whose
// full tiles do not fit (tetrahedra: about 70 vertex neighbours per cell).
std::vector<std::int32_t> split_tiles(const std::vector<std::int32_t>& tile_off, const Csr& dev_ccn, const Csr& dev_ccf,
                                      int max_slots);

}  // namespace cfdg
