// The solver's mesh container: the HDF5 layout of the reference's
// export_mesh_hdf5 / import_mesh_hdf5 (src/io/solver_mesh/hdf5_writer.cpp:195-389,
// hdf5_reader.cpp:140-461; tree in the reference README.md:394-462), for a
// single partition (nprocs = 1).
//
// Layout facts kept from the reference code (they differ from the README tree
// in places): the metadata are attributes of the root group; the partition
// tables are uint64 (nprocs + 1 entries) and int (part2rank / rank2part);
// cells/faces nodes_offsets hold one entry per cell / face (the trailing
// offset is not stored, the reader restores it from the node count);
// centroids, normals and coordinates are [n, 3] doubles; patch names and CGNS
// types are fixed 64-byte C strings; patch_face_offsets holds n_patches + 1
// entries per partition.  Local indices are int32; global ids are written as
// int32 when they fit (the reference's default GlobalIndex) and int64
// otherwise, and read as either.
//
// libhdf5 is loaded at run time (dlopen), so the library itself carries no
// link dependency on it: CFDG_HDF5_LIB names the library, else
// /opt/conda/lib/libhdf5.so.103, libhdf5.so.103 and libhdf5.so are tried.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace cfdg::front {

struct ContainerError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// MeshPart fields of one rank (include/cfd/mesh/localmesh.hpp:17-102)
struct ContainerMesh {
    std::uint64_t n_cells_g = 0, n_faces_g = 0, n_bfaces_g = 0, n_nodes_g = 0;
    std::array<double, 3> bbox_lo{0, 0, 0}, bbox_hi{0, 0, 0};

    std::vector<std::uint8_t> cell_type;
    std::vector<std::int64_t> cell_gid;
    std::vector<std::int32_t> cell_donor;  // -1 for owned cells
    std::vector<double> cell_volume, cell_cx, cell_cy, cell_cz;
    std::vector<std::int32_t> cell_nodes_off;  // n_cells + 1
    std::vector<std::int32_t> cell_nodes;

    std::vector<std::int32_t> face_owner, face_neigh, face_patch;  // neigh / patch -1 where absent
    std::vector<std::uint8_t> face_type;
    std::vector<double> face_area, face_nx, face_ny, face_nz, face_cx, face_cy, face_cz;
    std::vector<std::int32_t> face_nodes_off;  // n_faces + 1
    std::vector<std::int32_t> face_nodes;

    std::vector<std::int64_t> node_gid;
    std::vector<double> node_x, node_y, node_z;

    std::vector<std::int32_t> nb_ranks, send_offsets, recv_offsets, send_owned, recv_ghost;

    std::vector<std::string> patch_names, patch_cgns_types;
    std::vector<std::int32_t> patch_face_offsets;  // n_patches + 1
    std::vector<std::int32_t> patch_faces;

    // derived on read (hdf5_reader.cpp:298-307, 421-450)
    std::int32_t n_own = 0, n_internal_faces = 0, n_inner_faces = 0, n_nodes_own = 0;

    std::int32_t n_cells() const { return static_cast<std::int32_t>(cell_type.size()); }
    std::int32_t n_faces() const { return static_cast<std::int32_t>(face_owner.size()); }
    std::int32_t n_nodes() const { return static_cast<std::int32_t>(node_x.size()); }
};

// Reads the partition of rank 0 of a one-rank run; a file written for more
// ranks is rejected with the reference's message.  Throws ContainerError.
ContainerMesh read_container(const std::string& path);
// Writes a one-partition file (rank2part = part2rank = [0], empty /comm).
void write_container(const std::string& path, const ContainerMesh& m);

}  // namespace cfdg::front
