// Kernel argument blocks and launch wrappers shared by the host context code
// and the HIP translation units.  All device pointers; passed by value.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_physics.hpp"

namespace cfdg {

constexpr int kMaxPatches = 64;

// Mesh in device memory: face arrays in the 3-zone order, cell-centric CSR lists.
struct MeshArgs {
    int nc;       // cells
    int nf;       // faces
    int ninner;   // faces with a neighbour cell
    int ld;       // plane stride of the field arrays (n_total = nc + boundary faces)
    const int* owner;
    const int* neigh;
    const double* fnx;
    const double* fny;
    const double* fnz;
    const double* farea;
    const double* fcx;
    const double* fcy;
    const double* fcz;
    const double* dvx;   // owner -> neighbour (or mirrored ghost) centroid vector
    const double* dvy;
    const double* dvz;
    const double* dinv;  // 1 / |d|
    const double* ccx;
    const double* ccy;
    const double* ccz;
    const double* vol;
    const int* cf_off;   // cell -> faces, ascending face index
    const int* cf;
    const int* ccf_off;  // cell -> face neighbours, ascending cell index
    const int* ccf;
    const int* ccn_off;  // cell -> vertex neighbours (LSQ stencil), ascending
    const int* ccn;
    const double* lsq;   // LSQ coefficients, three planes of n_edges
    std::int64_t n_edges;
    const int* bface_bc; // boundary face -> patch id
};

struct FieldArgs {
    const double* U;  // conservative, 5 planes
    double* q;        // primitive (p, u, v, w, T), 5 planes incl. ghosts
    double* grad;     // 15 planes: (3 v + d)
    double* phi;      // 5 planes (cells only)
    double* res;      // 5 planes
    double* lam;      // spectral radius per cell
};

struct NumArgs {
    Gas gas;
    int viscous;
    int muscl;
    int limiter;
    int gradient;  // 0 GG face, 1 GG cell, 2 LSQ node
    double k3;     // venkat_k^3
};

struct BcTable {
    BcDev bc[kMaxPatches];
};

struct BsrArgs {
    int n;
    const int* row_ptr;
    const int* cols;
    const int* diag;
};

// --- residual stages (kernels_residual.hip) ---------------------------------
void launch_primitives(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, hipStream_t s);
void launch_ghosts(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, const BcTable& bcs, hipStream_t s);
void launch_gradient_limiter(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, hipStream_t s);
void launch_ghost_gradients(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, const BcTable& bcs,
                            hipStream_t s);
void launch_flux(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, hipStream_t s);
void launch_dt(const MeshArgs& m, const double* lam, double cfl, double* dt, double* alpha, hipStream_t s);
void launch_jacobian(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, const BsrArgs& b,
                     const double* alpha, double* values, hipStream_t s);

}  // namespace cfdg
