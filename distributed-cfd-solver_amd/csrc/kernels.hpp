// Kernel argument blocks, device data layouts and launch wrappers shared by the
// host context code and the HIP translation units.  All pointers are device
// pointers; blocks are passed by value.
//
// Device cell numbering: cells are renumbered into compact tiles (setup.hpp,
// tile_cells); faces keep their mesh index, so every per-cell sum over faces
// runs in ascending face index exactly as the reference's face sweeps do.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_physics.hpp"
#include "precision.hpp"

namespace cfdg {

constexpr int kMaxPatches = 64;
constexpr int kTileCells = 256;       // cells per tile (= threads per flux workgroup)
constexpr int kTileFaces = 1024;      // faces per tile held in LDS by the flux kernel
constexpr int kFluxThreads = 512;     // threads per flux workgroup (kTileFaces / kFluxThreads faces per lane)
constexpr int kNumXcd = 8;            // MI355X: 8 accelerator complex dies, one L2 each
constexpr int kGradSlots = 640;       // LDS slots (tile cells + stencil halo) of the staged gradient kernel
constexpr int kGradSlotsBig = 1088;   // its wide form, for large stencils (tetrahedra): 2 workgroups per CU
// storage of the Jacobian blocks (values of the precision enumeration of cfdg.h)
constexpr int kStoreF64 = 0, kStoreF32 = 1, kStoreBF16 = 2;

// Cell data: structure of arrays in the device (tile) order, kCellPlanes
// planes of FieldArgs::ld doubles (ld >= cells + boundary ghosts, a multiple of
// 64): plane k holds quantity k of every cell and ghost.  Planes are stored in
// interleaved pairs -- quantity k of cell c at ((k / 2) ld + c) 2 + k % 2 --
// so that one 16-byte load brings two quantities of a cell.  Consecutive lanes
// of a wave work on neighbouring cells of one tile, so every load of a wave
// falls in a few cache lines.
constexpr int kCellPlanes = 36;  // 35 used
constexpr int kRQ = 0;     // p, u, v, w, T
constexpr int kRC = 5;     // centroid x, y, z
constexpr int kRPhi = 8;   // limiter of the 5 primitives
constexpr int kRVol = 13;  // volume
constexpr int kRG = 14;    // gradients, (3 v + d)
constexpr int kRInv = 30;  // inverse LSQ normal matrix (00, 01, 02, 11, 12, 22), set-up time

// Face data: kFacePlanes planes of MeshArgs::fld doubles in mesh face order
// (unit normal, area, centroid, owner->neighbour centroid vector -- mirrored
// 2 (x_f - x_c) at boundary faces -- and its inverse length), plus
// (owner, neighbour) device ids per face, the neighbour of a boundary face
// being its ghost n_cells + (f - n_inner).  The flux kernel reads a second,
// shorter copy laid out in the order of the tiles' face lists: the first
// kTilePlanes - 1 planes (normal, area, centroid) and, in plane kTCells, the
// face's (owner, neighbour or ghost) ids packed into the 64 bits of the
// double (owner in the low half); it forms the centroid vector itself.
constexpr int kFacePlanes = 12;  // 11 used, paired like the cell planes
constexpr int kFN = 0, kFA = 3, kFC = 4, kFD = 7, kFDinv = 10;
constexpr int kTilePlanes = 8, kTCells = 7;
static_assert(kTCells % 2 == 1 && kTCells == kFC + 3, "packed ids share a pair with the centroid z");

// Position of quantity k of entry i in paired planes of stride ld.
__host__ __device__ __forceinline__ long long plane_at(int k, long long ld, long long i) {
    return ((k >> 1) * ld + i) * 2 + (k & 1);
}

// Read / write views of one cell or face in the planes.
struct CellRef {
    const double* b;
    long long ld;
    int c;
    // uniform plane-pair base + 32-bit per-lane offset (scalar base address + vector offset loads)
    __host__ __device__ __forceinline__ double operator[](int k) const {
        const double* base = b + static_cast<long long>(k >> 1) * 2 * ld;
        return base[2u * static_cast<unsigned>(c) + static_cast<unsigned>(k & 1)];
    }
};
struct CellRefW {
    double* b;
    long long ld;
    int c;
    __host__ __device__ __forceinline__ double& operator[](int k) const { return b[plane_at(k, ld, c)]; }
};
struct FaceRef {
    const double* b;
    long long ld;
    int f;
    __host__ __device__ __forceinline__ double operator[](int k) const { return b[plane_at(k, ld, f)]; }
};

// Face planes read once per launch (the tile face-list stream): non-temporal
// loads, so the stream does not evict the cell planes that faces share.
struct FaceRefNT {
    const double* b;
    long long ld;
    int f;
    __device__ __forceinline__ double operator[](int k) const {
        const double* base = b + static_cast<long long>(k >> 1) * 2 * ld;
        return __builtin_nontemporal_load(base + (2u * static_cast<unsigned>(f) + static_cast<unsigned>(k & 1)));
    }
};

struct MeshArgs {
    int nc;        // cells
    int nf;        // faces
    int ninner;    // faces with a neighbour cell
    int ntiles;
    const double* fp;         // kFacePlanes x fld, mesh face order
    long long fld;
    const int2* fcells;       // (owner, neighbour or ghost) per face, device ids
    const double* tf;         // kTilePlanes x tld, tile face-list order
    long long tld;
    const int* cf_off;        // device cell -> faces, ascending face index
    const int* cf;
    const int* cf_slot;       // position of the face in the tile's face list (negative: neighbour side)
    const int* ccf_off;       // device cell -> face neighbours (device ids, ascending mesh index)
    const int* ccf;
    const int* ccn_off;       // device cell -> vertex neighbours (LSQ stencil), ascending mesh index
    const int* ccn;
    const int* bface_bc;      // boundary face -> patch id
    const int* tile_off;      // tiles + 1
    const int* tile_face_off; // tiles + 1
    // LDS-staged gradient stencils over the gradient tiles (null when they do not fit)
    const int* gtile_off;     // gradient tiles + 1: the flux tiles, or consecutive sub-ranges of them
    int ngtiles;
    int gblock;               // threads per gradient workgroup (multiple of 64, >= cells of any gradient tile)
    int gslots;               // LDS slot capacity of the gradient kernel: kGradSlots or kGradSlotsBig
    const int* tile_halo_off; // gradient tiles + 1
    const int* tile_halo;     // halo cells of each gradient tile, ascending device id
    // LDS slot of each ccn / ccf entry, per gradient tile in chunks of four:
    // entry k of the tile's i-th cell at ccn_base[t] + ((k/4) nt + i) 4 + k%4
    // (nt = cells of tile t)
    const std::uint16_t* ccn_slot;
    const std::uint16_t* ccf_slot;
    const int* ccn_base;      // gradient tiles + 1
    const int* ccf_base;
    // x_f - x_c of every cell face in ascending face index, per gradient tile:
    // component d of face k of the tile's i-th cell at
    // face_vec_base[t] + (d W + k) nt + i, W = the tile's longest face list
    const double* face_vec;
    const long long* face_vec_base;  // gradient tiles + 1
};

struct FieldArgs {
    const double* U;  // conservative, 5 planes of nc (device order)
    double* cs;       // cell planes, kCellPlanes x ld
    long long ld;
    double* res;      // 5 planes of nc
    double* lam;      // nc
    double* bflux;    // F (inviscid + viscous) of every boundary face, 5 per face
    const double* mut;  // eddy viscosity (cells + ghosts, ghosts zero), null without turbulence
    double* mdot;       // mass flux F[0] of every face (mesh order), gathered from mdot_tile on request
    double* mdot_tile;  // mass flux per tile face-list entry (tile stream order), written when mut is set
};

// Zone treatment of nu_tilde at a boundary face (spalart_allmaras.hpp:54-59)
constexpr int kZoneWall = 0, kZoneSymmetry = 1, kZoneInflow = 2, kZoneOutflow = 3;

// Spalart-Allmaras module fields, device order; arrays of ntot = cells +
// boundary ghosts hold the ghost of boundary face f at n_cells + (f - n_inner).
struct SaArgs {
    double nu_inf;        // freestream nu_tilde
    const int* kind;      // zone of each boundary face
    const double* dist;   // wall distance (nc)
    double* rnu;          // rho nu_tilde (ntot): cells = state, ghosts = zone values
    double* nu;           // nu_tilde (ntot)
    double* rho;          // density (ntot)
    double* mut;          // eddy viscosity (ntot, ghosts zero)
    double* g;            // nu_tilde gradient, 3 planes of nc
    double* res;          // module residual (nc)
    double* diag;         // point-implicit diagonal (nc)
    const double* mdot;   // mean-flow face mass flux (nf, mesh order)
    const double* mdot_tile;  // the same per tile face-list entry
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
    // per entry of the cell -> face lists (MeshArgs::cf_off order): the block
    // (row = the cell, column = the face's other cell) of an interior face, -1
    // at a boundary face
    const int* cf_blk;
};

// Bijective block -> work-item map that gives each XCD one contiguous range of
// the work (workgroups are dealt to the 8 XCDs round-robin), so cells that
// share neighbours are processed by the same L2.
__host__ __device__ __forceinline__ int xcd_block(int b, int nb) {
    const int x = b % kNumXcd, i = b / kNumXcd;
    const int base = nb / kNumXcd, rem = nb % kNumXcd;
    return x * base + (x < rem ? x : rem) + i;
}

// --- residual stages (kernels_residual.hip) ---------------------------------
void launch_primitives(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, hipStream_t s);
void launch_ghosts(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, const BcTable& bcs, hipStream_t s);
// With sa given and the LDS-staged kernel in use, the module's nu_tilde gradient
// is formed in the same pass; returns whether it was.
bool launch_gradient_limiter(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, hipStream_t s,
                             const SaArgs* sa = nullptr);
// inverse LSQ normal matrices (planes kRInv..), once the centroids and the vertex stencil are on the device
void launch_lsq_setup(const MeshArgs& m, const FieldArgs& f, hipStream_t s);
void launch_ghost_gradients(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, const BcTable& bcs,
                            hipStream_t s);
void launch_flux(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, hipStream_t s);
// out[5 p + v] = sum of F[v] over the boundary faces of patch p, p < n_patches
void launch_boundary_integrals(const MeshArgs& m, const FieldArgs& f, int n_patches, double* out, hipStream_t s);
// unsteady = c0 / dt of dual time stepping (0 in steady mode)
void launch_dt(const MeshArgs& m, const FieldArgs& f, const double* lam, double cfl, double unsteady, double* dt,
               double* alpha, hipStream_t s);
// rho, u, v, w, p, Mach of every cell (6 planes of nc, device order) for the solution files
void launch_output_fields(int nc, const Gas& g, const double* U, double* out, hipStream_t s);
// mdot[f] = mdot_tile[face_entry[f]] for every face (mesh order)
void launch_gather_mdot(int nf, const int* face_entry, const double* mdot_tile, double* mdot, hipStream_t s);
// forward Euler: alpha = CFL / lambda, then U <- U - alpha R
void launch_dt_explicit(const MeshArgs& m, const FieldArgs& f, const double* lam, double cfl, double* dt,
                        double* alpha, hipStream_t s);
void launch_explicit_update(const MeshArgs& m, const FieldArgs& f, const double* alpha, double* U, hipStream_t s);
// dual time stepping: R += V (c0 U + c1 U^n [+ c2 U^{n-1}]), bdf = 1 or 2
void launch_unsteady_source(const MeshArgs& m, const FieldArgs& f, int bdf, double inv_dt, const double* Un,
                            const double* Unm1, hipStream_t s);
// values: nnzb x 25 entries of the storage type (kStoreF64 / kStoreF32 / kStoreBF16)
void launch_jacobian(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, const BsrArgs& b,
                     const double* alpha, void* values, int storage, hipStream_t s);

// --- Spalart-Allmaras module (kernels_sa.hip) --------------------------------
void launch_sa_prepare(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, const SaArgs& sa, hipStream_t s);
void launch_sa_gradient(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, const SaArgs& sa, hipStream_t s);
// res / diag of the module and lam += the interior faces' module spectral radius
void launch_sa_residual(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, const SaArgs& sa, hipStream_t s);
void launch_sa_update(int nc, int point_implicit, const double* alpha, const SaArgs& sa, hipStream_t s);

}  // namespace cfdg
