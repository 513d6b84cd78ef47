// Mesh preprocessing on the device: the cell and face metrics of a mesh
// (src/mesh/geometry.cpp:301-539) from its nodes and node lists, one thread
// per cell and one per face, with the arithmetic of geometry.hpp that the
// host's compute_geometry runs too.  Every output element has one writer.
#include <hip/hip_runtime.h>

#include "geometry.hpp"
#include "kernels_mesh.hpp"
#include "mesh.hpp"

namespace cfdg {

namespace {

constexpr int kBlock = 256;

// Cells have 4 (tetrahedron) or 8 (hexahedron) nodes; the host checks this
// before the launch.
__global__ void __launch_bounds__(kBlock) k_cell_geometry(GeomArgs a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nc) return;
    double bx[8], by[8], bz[8];
    const int b = a.cell_off[c], e = a.cell_off[c + 1];
    for (int k = 0; k < 8 && b + k < e; ++k) {
        const int n = a.cell_nodes[b + k];
        bx[k] = a.x[n];
        by[k] = a.y[n];
        bz[k] = a.z[n];
    }
    const int type = (a.cell_type[c] == static_cast<std::uint8_t>(CellKind::Tet)) ? 0 : 3;
    const geom::CellMetrics cm = geom::cell_metrics(type, bx, by, bz);
    a.ccx[c] = cm.cx;
    a.ccy[c] = cm.cy;
    a.ccz[c] = cm.cz;
    a.cvol[c] = cm.vol;
}

// Faces have 3 to 8 nodes; the host checks this before the launch.
__global__ void __launch_bounds__(kBlock) k_face_geometry(GeomArgs a) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.nf) return;
    double bx[8], by[8], bz[8];
    const int b = a.face_off[f], e = a.face_off[f + 1];
    const int nn = e - b;
    for (int k = 0; k < 8 && k < nn; ++k) {
        const int n = a.face_nodes[b + k];
        bx[k] = a.x[n];
        by[k] = a.y[n];
        bz[k] = a.z[n];
    }
    const geom::FaceMetrics fm = geom::face_metrics(nn, bx, by, bz);
    a.fcx[f] = fm.cx;
    a.fcy[f] = fm.cy;
    a.fcz[f] = fm.cz;
    a.fnx[f] = fm.nx;
    a.fny[f] = fm.ny;
    a.fnz[f] = fm.nz;
    a.farea[f] = fm.area;
}

}  // namespace

void launch_mesh_geometry(const GeomArgs& a, void* stream) {
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (a.nc > 0) hipLaunchKernelGGL(k_cell_geometry, dim3((a.nc + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a);
    if (a.nf > 0) hipLaunchKernelGGL(k_face_geometry, dim3((a.nf + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a);
}

}  // namespace cfdg
