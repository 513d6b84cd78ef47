// Mesh preprocessing on the device (kernels_mesh.hip).
#pragma once

#include <cstdint>

namespace cfdg {

// Device arrays of one mesh's geometry computation: nodes (3 x n_nodes),
// cell and face node lists (CSR), cell kinds (CellKind), and the outputs:
// cell centroid (3 x nc) and volume, face centroid (3 x nf), unit normal
// (3 x nf) and area.
struct GeomArgs {
    int nc, nf;
    const double *x, *y, *z;
    const int *cell_off, *cell_nodes;
    const std::uint8_t* cell_type;
    const int *face_off, *face_nodes;
    double *ccx, *ccy, *ccz, *cvol;
    double *fcx, *fcy, *fcz, *fnx, *fny, *fnz, *farea;
};

// One thread per cell and one per face on the stream s (geometry.cpp:301-539).
void launch_mesh_geometry(const GeomArgs& a, void* stream);

}  // namespace cfdg
