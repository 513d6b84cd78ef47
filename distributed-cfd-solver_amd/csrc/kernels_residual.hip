// Residual, time-step and Jacobian kernels (gfx950).
//
// Data: cell and face planes (structure of arrays, kernels.hpp).  Cells are
// in the device (tile) order; faces keep their mesh index.
//
//   k_primitives        U -> (p,u,v,w,T) into the cell records
//   k_ghosts            boundary ghost states, one thread per boundary face
//   k_gradient_limiter  cell gradient (node-stencil least squares with the
//                       3x3 normal equations rebuilt per cell from neighbour
//                       centroids, or Green-Gauss) + MUSCL limiter, one thread
//                       per cell; k_gradient_limiter_lds is the same per cell
//                       with the tile's neighbour stencil staged in LDS
//   k_ghost_gradients   boundary ghost gradients
//   k_flux              tiled flux residual: a workgroup owns a tile of cells,
//                       evaluates each face touching the tile once into LDS,
//                       then every cell sums its faces in ascending face index
//   k_dt, k_jacobian    local time step; Picard Jacobian on the flux tiles, one
//                       thread per block row
//
// No floating-point atomics: every output element has a single writer and
// every sum has a fixed order.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace cfdg {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int face_owner(const MeshArgs& m, int f) { return m.fcells[f].x; }
// neighbour cell, or the ghost n_cells + (f - n_inner) of a boundary face
__device__ __forceinline__ int face_other(const MeshArgs& m, int f) { return m.fcells[f].y; }
__device__ __forceinline__ FaceRef frec(const MeshArgs& m, int f) { return FaceRef{m.fp, m.fld, f}; }
__device__ __forceinline__ CellRef crec(const FieldArgs& f, int c) { return CellRef{f.cs, f.ld, c}; }
__device__ __forceinline__ CellRefW crecw(const FieldArgs& f, int c) { return CellRefW{f.cs, f.ld, c}; }

// U -> q for the cells (solver.hpp:185-188)
__global__ void __launch_bounds__(kBlock) k_primitives(MeshArgs m, FieldArgs f, NumArgs nm) {
    const int c = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (c >= m.nc) return;
    double U[5], q[5];
    for (int v = 0; v < 5; ++v) U[v] = f.U[static_cast<long long>(v) * m.nc + c];
    cons_to_prim(nm.gas, U, q);
    const CellRefW r = crecw(f, c);
    for (int v = 0; v < 5; ++v) r[kRQ + v] = q[v];
}

// 1 / max((x_f - x_c) . n, 1e-14)   (boundary_traversal.hpp:98-108)
template <class FR, class CR>
__device__ __forceinline__ double rcfn_inverse(const FR& fr, const CR& cr) {
    const double rx = fr[kFC] - cr[kRC], ry = fr[kFC + 1] - cr[kRC + 1], rz = fr[kFC + 2] - cr[kRC + 2];
    const double r = rx * fr[kFN] + ry * fr[kFN + 1] + rz * fr[kFN + 2];
    return 1.0 / dmax(r, 1.0e-14);
}

// Ghost primitive states (bc/physical/*.hpp value kernels)
__global__ void __launch_bounds__(kBlock) k_ghosts(MeshArgs m, FieldArgs f, NumArgs nm, BcTable bt) {
    const int bf = blockIdx.x * blockDim.x + threadIdx.x;
    const int nbf = m.nf - m.ninner;
    if (bf >= nbf) return;
    const int fc = m.ninner + bf;
    const FaceRef fr = frec(m, fc);
    const CellRef ri = crec(f, face_owner(m, fc));
    const CellRefW rg = crecw(f, m.nc + bf);
    const BcDev bp = bt.bc[m.bface_bc[bf]];
    double qi[5], qg[5];
    for (int v = 0; v < 5; ++v) qi[v] = ri[kRQ + v];
    const double nx = fr[kFN], ny = fr[kFN + 1], nz = fr[kFN + 2];
    switch (bp.type) {
        case kSupIn: {
            const double qf[5] = {bp.p, bp.u, bp.v, bp.w, bp.T};
            for (int v = 0; v < 5; ++v) qg[v] = 2.0 * qf[v] - qi[v];
            qg[0] = dmax(qg[0], 1.0);
            qg[4] = dmax(qg[4], 1.0);
            break;
        }
        case kSlip:
        case kSym: {
            const double vn2 = 2.0 * (qi[1] * nx + qi[2] * ny + qi[3] * nz);
            qg[1] = qi[1] - vn2 * nx;
            qg[2] = qi[2] - vn2 * ny;
            qg[3] = qi[3] - vn2 * nz;
            qg[0] = qi[0];
            qg[4] = qi[4];
            break;
        }
        case kFar: {
            double qb[5];
            bool so = false;
            farfield_state(bp, nm.gas.gamma, nm.gas.R, qi[0], qi[4], qi[1], qi[2], qi[3], nx, ny, nz, qb, so);
            if (so) {
                for (int v = 0; v < 5; ++v) qg[v] = qi[v];
            } else {
                for (int v = 0; v < 5; ++v) qg[v] = 2.0 * qb[v] - qi[v];
                qg[0] = dmax(qg[0], 1.0);
                qg[4] = dmax(qg[4], 1.0);
            }
            break;
        }
        case kNoSlip:
        case kSubIn:
            qg[0] = qi[0];
            qg[1] = 2.0 * bp.u - qi[1];
            qg[2] = 2.0 * bp.v - qi[2];
            qg[3] = 2.0 * bp.w - qi[3];
            qg[4] = 2.0 * bp.T - qi[4];
            qg[4] = dmax(qg[4], 1.0);
            break;
        case kNoSlipHF: {
            const double rinv = rcfn_inverse(fr, ri);
            qg[0] = qi[0];
            qg[1] = 2.0 * bp.u - qi[1];
            qg[2] = 2.0 * bp.v - qi[2];
            qg[3] = 2.0 * bp.w - qi[3];
            const double r = 1.0 / rinv;
            qg[4] = qi[4] + 2.0 * bp.tmp_grad * r;
            qg[4] = dmax(qg[4], 1.0);
            break;
        }
        case kSubOut:
            qg[0] = 2.0 * bp.p - qi[0];
            qg[0] = dmax(qg[0], 1.0);
            for (int v = 1; v < 5; ++v) qg[v] = qi[v];
            break;
        default:  // supersonic outlet: extrapolation
            for (int v = 0; v < 5; ++v) qg[v] = qi[v];
            break;
    }
    for (int v = 0; v < 5; ++v) rg[kRQ + v] = qg[v];
}

// Inverse of the weighted least-squares normal matrix of cell c over its
// vertex-sharing stencil (LeastSquaresCellNodeGradient::setup,
// gradient_operator.cpp:622-707): A = sum w dr dr^T + 1e-14 I with
// w = 1/max(|dr|^2, 1e-30), inverted by the adjugate (zero when |det| <= 1e-14).
// A is symmetric and so is its adjugate inverse entry for entry, so six values
// (00, 01, 02, 11, 12, 22) hold it.
__device__ __forceinline__ void lsq_inverse(const MeshArgs& m, const FieldArgs& f, int c, double out[6]) {
    const CellRef rc = crec(f, c);
    const double x0 = rc[kRC], y0 = rc[kRC + 1], z0 = rc[kRC + 2];
    double A0 = 0.0, A1 = 0.0, A2 = 0.0, A4 = 0.0, A5 = 0.0, A8 = 0.0;
    for (int j = m.ccn_off[c]; j < m.ccn_off[c + 1]; ++j) {
        const CellRef rn = crec(f, m.ccn[j]);
        const double dx = rn[kRC] - x0, dy = rn[kRC + 1] - y0, dz = rn[kRC + 2] - z0;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double w = 1.0 / dmax(r2, 1.0e-30);
        A0 += w * dx * dx;
        A1 += w * dx * dy;
        A2 += w * dx * dz;
        A4 += w * dy * dy;
        A5 += w * dy * dz;
        A8 += w * dz * dz;
    }
    const double A3 = A1, A6 = A2, A7 = A5;
    A0 += 1.0e-14;
    A4 += 1.0e-14;
    A8 += 1.0e-14;
    for (int k = 0; k < 6; ++k) out[k] = 0.0;
    const double det = A0 * (A4 * A8 - A5 * A7) - A1 * (A3 * A8 - A5 * A6) + A2 * (A3 * A7 - A4 * A6);
    if (!(fabs(det) <= 1.0e-14)) {
        const double id = 1.0 / det;
        out[0] = (A4 * A8 - A5 * A7) * id;
        out[1] = (A2 * A7 - A1 * A8) * id;
        out[2] = (A1 * A5 - A2 * A4) * id;
        out[3] = (A0 * A8 - A2 * A6) * id;
        out[4] = (A2 * A3 - A0 * A5) * id;
        out[5] = (A0 * A4 - A1 * A3) * id;
    }
}

// Set-up: the inverse normal matrix of every cell into planes kRInv..kRInv+5
// (centroids only, so once per mesh).
__global__ void __launch_bounds__(kBlock) k_lsq_setup(MeshArgs m, FieldArgs f) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nc) return;
    double inv[6];
    lsq_inverse(m, f, c, inv);
    const CellRefW w = crecw(f, c);
    for (int k = 0; k < 6; ++k) w[kRInv + k] = inv[k];
}

// Cell gradient + MUSCL limiter of device cell c (gradient_operator.cpp:549-707,
// :197-278; muscl.hpp:26-104).  `self`, nbn(j) and nbf(j) point at 8 doubles
// laid out as the first half of a cell record (q, centroid) for the cell, its
// j-th vertex neighbour and its j-th face neighbour -- either the records in
// planes in HBM or their copies staged in LDS.
// Optional nu_tilde gradient formed in the same least-squares pass (the
// turbulence module's scalar, same coefficients, same order per term).
struct NoScalar {
    static constexpr bool kOn = false;
    __device__ double self() const { return 0.0; }
    template <class Nb>
    __device__ double nb(const Nb&) const { return 0.0; }
    __device__ void store(int, const double*) const {}
};
struct LdsScalar {
    static constexpr bool kOn = true;
    const double* s;       // LDS slots (the neighbour's slot comes with its view)
    int own;               // slot of the cell
    double* g;             // output: 3 planes of nc
    long long nc;
    __device__ double self() const { return s[own]; }
    template <class Nb>
    __device__ double nb(const Nb& rn) const { return s[rn.s]; }
    __device__ void store(int c, const double* v) const {
        g[c] = v[0];
        g[nc + c] = v[1];
        g[2 * nc + c] = v[2];
    }
};

template <class Self, class NbF, class FVec>
__device__ __forceinline__ void store_and_limit(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, int c,
                                                const Self& self, const double (&g)[5][3], NbF nbf, FVec fvec);

// fvec(e, k, x0, y0, z0, ex, ey, ez) gives x_f - x_c of the cell's k-th face
// (entry e of cf).
template <class Self, class NbN, class NbF, class FVec, class Sc = NoScalar>
__device__ __forceinline__ void gradient_limiter_cell(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, int c,
                                                      const Self& self, NbN nbn, NbF nbf, FVec fvec,
                                                      const Sc& sc = Sc{}) {
    const CellRefW rc = crecw(f, c);
    double qc[5];
    for (int v = 0; v < 5; ++v) qc[v] = self[kRQ + v];
    const double x0 = self[kRC], y0 = self[kRC + 1], z0 = self[kRC + 2];
    double g[5][3];
    for (int v = 0; v < 5; ++v) g[v][0] = g[v][1] = g[v][2] = 0.0;

    if (nm.gradient == 2) {
        // weighted least squares over the vertex-sharing stencil with the
        // inverse normal matrix of the cell (set-up time, lsq_inverse): the
        // coefficients C_j = A^-1 w_j dr_j are applied on the fly
        const int jb = m.ccn_off[c], je = m.ccn_off[c + 1];
        double gs[3] = {0.0, 0.0, 0.0};
        const double sc0 = sc.self();
        if (je > jb) {
            const CellRef ri{f.cs, f.ld, c};
            const double i0 = ri[kRInv], i1 = ri[kRInv + 1], i2 = ri[kRInv + 2], i4 = ri[kRInv + 3],
                         i5 = ri[kRInv + 4], i8 = ri[kRInv + 5];
            for (int j = jb, k = 0; j < je; ++j, ++k) {
                const auto rn = nbn(j, k);
                const double dx = rn[kRC] - x0, dy = rn[kRC + 1] - y0, dz = rn[kRC + 2] - z0;
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double w = 1.0 / dmax(r2, 1.0e-30);
                const double rx = w * dx, ry = w * dy, rz = w * dz;
                const double cx = i0 * rx + i1 * ry + i2 * rz;
                const double cy = i1 * rx + i4 * ry + i5 * rz;
                const double cz = i2 * rx + i5 * ry + i8 * rz;
                for (int v = 0; v < 5; ++v) {
                    const double d = rn[kRQ + v] - qc[v];
                    g[v][0] += cx * d;
                    g[v][1] += cy * d;
                    g[v][2] += cz * d;
                }
                if constexpr (Sc::kOn) {
                    const double d = sc.nb(rn) - sc0;
                    gs[0] += cx * d;
                    gs[1] += cy * d;
                    gs[2] += cz * d;
                }
            }
        }
        if constexpr (Sc::kOn) sc.store(c, gs);
    } else {
        // Green-Gauss: the face scatter gathered per cell (both the face- and
        // cell-based variants form these sums)
        for (int e = m.cf_off[c]; e < m.cf_off[c + 1]; ++e) {
            const int fc = m.cf[e];
            const FaceRef fr = frec(m, fc);
            const int o = face_owner(m, fc);
            const int nb = face_other(m, fc);
            const double a05 = 0.5 * fr[kFA];
            const double hx = a05 * fr[kFN], hy = a05 * fr[kFN + 1], hz = a05 * fr[kFN + 2];
            const bool own = (o == c);
            const CellRef ro = crec(f, o);
            const CellRef rn = crec(f, nb);
            for (int v = 0; v < 5; ++v) {
                const double s = ro[kRQ + v] + rn[kRQ + v];
                const double fx = hx * s, fy = hy * s, fz = hz * s;
                if (own) {
                    g[v][0] += fx;
                    g[v][1] += fy;
                    g[v][2] += fz;
                } else {
                    g[v][0] -= fx;
                    g[v][1] -= fy;
                    g[v][2] -= fz;
                }
            }
        }
        const double iv = 1.0 / rc[kRVol];
        for (int v = 0; v < 5; ++v) {
            g[v][0] *= iv;
            g[v][1] *= iv;
            g[v][2] *= iv;
        }
    }
    store_and_limit(m, f, nm, c, self, g, nbf, fvec);
}

// The gradients g of device cell c stored, then the MUSCL limiter of the cell
// from them (muscl.hpp:26-104): the second half of gradient_limiter_cell.
template <class Self, class NbF, class FVec>
__device__ __forceinline__ void store_and_limit(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, int c,
                                                const Self& self, const double (&g)[5][3], NbF nbf, FVec fvec) {
    const CellRefW rc = crecw(f, c);
    double qc[5];
    for (int v = 0; v < 5; ++v) qc[v] = self[kRQ + v];
    const double x0 = self[kRC], y0 = self[kRC + 1], z0 = self[kRC + 2];
    for (int v = 0; v < 5; ++v)
        for (int d = 0; d < 3; ++d) rc[kRG + 3 * v + d] = g[v][d];

    if (!nm.muscl) return;
    double qmax[5], qmin[5];
    for (int v = 0; v < 5; ++v) qmax[v] = qmin[v] = qc[v];
    for (int j = m.ccf_off[c], k = 0; j < m.ccf_off[c + 1]; ++j, ++k) {
        const auto rn = nbf(j, k);
        for (int v = 0; v < 5; ++v) {
            const double qn = rn[kRQ + v];
            qmax[v] = dmax(qmax[v], qn);
            qmin[v] = dmin(qmin[v], qn);
        }
    }
    double ph[5] = {1.0, 1.0, 1.0, 1.0, 1.0};
    const double eps2 = nm.k3 * rc[kRVol];
    for (int e = m.cf_off[c], k = 0; e < m.cf_off[c + 1]; ++e, ++k) {
        double ex, ey, ez;
        fvec(e, k, x0, y0, z0, ex, ey, ez);
        for (int v = 0; v < 5; ++v) {
            const double df = g[v][0] * ex + g[v][1] * ey + g[v][2] * ez;
            const double dnb = (df > 0.0) ? (qmax[v] - qc[v]) : (qmin[v] - qc[v]);
            ph[v] = dmin(ph[v], limiter_phi(nm.limiter, dnb, df, eps2));
        }
    }
    for (int v = 0; v < 5; ++v) rc[kRPhi + v] = ph[v];
}

// One thread per cell, neighbours gathered from the records in HBM.
__global__ void __launch_bounds__(kBlock) k_gradient_limiter(MeshArgs m, FieldArgs f, NumArgs nm) {
    const int c = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (c >= m.nc) return;
    gradient_limiter_cell(
        m, f, nm, c, crec(f, c), [&](int j, int) { return crec(f, m.ccn[j]); },
        [&](int j, int) { return crec(f, m.ccf[j]); },
        [&](int e, int, double x0, double y0, double z0, double& ex, double& ey, double& ez) {
            const FaceRef fr = frec(m, m.cf[e]);
            ex = fr[kFC] - x0;
            ey = fr[kFC + 1] - y0;
            ez = fr[kFC + 2] - z0;
        });
}

// Four consecutive 16-bit LDS slot indices of a stencil list, loaded as one
// 8-byte word; entry k of the chunk sits in bits 16 (k mod 4) and up.  Taking
// an entry by shifting the word (not by picking one of two 32-bit halves)
// keeps the chunk in registers: a pick between the halves by a run-time index
// was compiled into an indexed load from a stack copy.
typedef unsigned long long SlotChunk;
__device__ __forceinline__ int chunk_slot(SlotChunk ch, int k) {
    return static_cast<int>((ch >> ((k & 3) * 16)) & 0xffffULL);
}

// View of one LDS slot of the staged gradient kernel: planes 0-7 (q,
// centroid) stored as interleaved pairs like the cell planes, quantity k of
// slot s at sq[((k / 2) kSlots + s) 2 + k % 2], so one 16-byte LDS read
// brings two quantities of a neighbour.
template <int kSlots>
struct LdsSlot {
    const double* sq;
    int s;
    __device__ __forceinline__ double operator[](int k) const { return sq[((k >> 1) * kSlots + s) * 2 + (k & 1)]; }
};

// One workgroup per gradient tile (a flux tile, or a consecutive sub-range of
// one where the tile's stencil halo would not fit; blockDim.x >= its cells):
// the (q, centroid) planes of the tile's cells and of its stencil halo are
// staged once into LDS (structure of arrays, so that lanes reading different
// slots of one plane spread over the banks), then every cell reads its vertex
// and face neighbours from there.
template <bool kSA, int kSlots>
__global__ void __launch_bounds__(kTileCells) k_gradient_limiter_lds(MeshArgs m, FieldArgs f, NumArgs nm, SaArgs sa) {
    __shared__ __attribute__((aligned(16))) double sq[(kSA ? 9 : 8) * kSlots];  // 4 plane pairs (+ nu_tilde)
    const int t = xcd_block(blockIdx.x, gridDim.x);
    const int c0 = m.gtile_off[t], nt = m.gtile_off[t + 1] - c0;
    const int hb = m.tile_halo_off[t], ns = nt + (m.tile_halo_off[t + 1] - hb);
    // slot lists in chunks of four entries (8 bytes) per cell, consecutive
    // lanes on consecutive chunks; the first two chunks of each list are
    // loaded before the staging so that their latency overlaps it
    const bool live = static_cast<int>(threadIdx.x) < nt;
    const int nchn = (m.ccn_base[t + 1] - m.ccn_base[t]) / (4 * nt);
    const int nchf = (m.ccf_base[t + 1] - m.ccf_base[t]) / (4 * nt);
    const SlotChunk* chn = reinterpret_cast<const SlotChunk*>(m.ccn_slot + m.ccn_base[t]) + threadIdx.x;
    const SlotChunk* chf = reinterpret_cast<const SlotChunk*>(m.ccf_slot + m.ccf_base[t]) + threadIdx.x;
    SlotChunk curn = 0, nxtn = 0, curf = 0, nxtf = 0;
    if (live) {
        if (nchn > 0) curn = chn[0];
        if (nchn > 1) nxtn = chn[nt];
        if (nchf > 0) curf = chf[0];
        if (nchf > 1) nxtf = chf[nt];
    }
    for (int sl = threadIdx.x; sl < ns; sl += blockDim.x) {
        const int src = (sl < nt) ? c0 + sl : m.tile_halo[hb + sl - nt];
        for (int k = 0; k < 8; k += 2) {  // planes 0-7 = q, centroid: four 16-byte pair loads
            const double2 v = *reinterpret_cast<const double2*>(f.cs + plane_at(k, f.ld, src));
            *reinterpret_cast<double2*>(sq + ((k >> 1) * kSlots + sl) * 2) = v;
        }
        if constexpr (kSA) sq[8 * kSlots + sl] = sa.nu[src];
    }
    __syncthreads();
    if (!live) return;
    // entries are taken in order k = 0, 1, 2, ...: at each chunk boundary the
    // held next chunk becomes current and the one after it is requested
    const auto take = [&](int k, SlotChunk& cur, SlotChunk& nxt, const SlotChunk* ch, int nch) {
        if ((k & 3) == 0 && k > 0) {
            cur = nxt;
            if ((k >> 2) + 1 < nch) nxt = ch[((k >> 2) + 1) * nt];
        }
        return chunk_slot(cur, k);
    };
    const auto nbn = [&](int, int k) { return LdsSlot<kSlots>{sq, take(k, curn, nxtn, chn, nchn)}; };
    const auto nbf = [&](int, int k) { return LdsSlot<kSlots>{sq, take(k, curf, nxtf, chf, nchf)}; };
    // set-up-time face vectors, transposed per tile like the slot lists
    const long long vb = m.face_vec_base[t];
    const long long vw = (m.face_vec_base[t + 1] - vb) / (3LL * nt);
    const double* fvp = m.face_vec + vb + threadIdx.x;
    const auto fvec = [&](int, int k, double, double, double, double& ex, double& ey, double& ez) {
        ex = fvp[k * static_cast<long long>(nt)];
        ey = fvp[(vw + k) * nt];
        ez = fvp[(2 * vw + k) * nt];
    };
    if constexpr (kSA) {
        // the module's nu_tilde gradient in the same pass (k_sa_gradient's arithmetic)
        const LdsScalar sc{sq + 8 * kSlots, static_cast<int>(threadIdx.x), sa.g, m.nc};
        gradient_limiter_cell(m, f, nm, c0 + threadIdx.x, LdsSlot<kSlots>{sq, static_cast<int>(threadIdx.x)}, nbn, nbf,
                              fvec, sc);
    } else {
        gradient_limiter_cell(m, f, nm, c0 + threadIdx.x, LdsSlot<kSlots>{sq, static_cast<int>(threadIdx.x)}, nbn, nbf,
                              fvec);
    }
}

// The wide staged form with two lanes per cell (workgroups of up to 512
// lanes): its 1,088 LDS slots allow two workgroups per CU, so one lane per
// cell would leave two waves per SIMD to hide the stencil loop's latencies.
// The two lanes of a cell (adjacent lanes) split the cell's vertex stencil by
// chunks of four entries (lane h takes chunks h, h + 2, ...), each adding
// its entries in stencil order with the arithmetic of gradient_limiter_cell;
// the second lane's sums are then added to the first lane's, which stores the
// gradients and forms the limiter.  The sums are therefore those of the
// one-lane form up to the order of that one addition.  Node LSQ only (the
// only gradient the staged forms serve).
template <bool kSA>
__global__ void __launch_bounds__(2 * kTileCells) k_gradient_limiter_pair(MeshArgs m, FieldArgs f, NumArgs nm,
                                                                          SaArgs sa) {
    constexpr int kSlots = kGradSlotsBig;
    __shared__ __attribute__((aligned(16))) double sq[(kSA ? 9 : 8) * kSlots];
    const int t = xcd_block(blockIdx.x, gridDim.x);
    const int c0 = m.gtile_off[t], nt = m.gtile_off[t + 1] - c0;
    const int hb = m.tile_halo_off[t], ns = nt + (m.tile_halo_off[t + 1] - hb);
    const int cell = static_cast<int>(threadIdx.x >> 1), half = static_cast<int>(threadIdx.x & 1);
    const bool live = cell < nt;
    const int nchn = (m.ccn_base[t + 1] - m.ccn_base[t]) / (4 * nt);
    const SlotChunk* chn = reinterpret_cast<const SlotChunk*>(m.ccn_slot + m.ccn_base[t]) + cell;
    // this lane's first two vertex-slot chunks, requested before the staging
    SlotChunk cur = 0, nxt = 0;
    if (live) {
        if (half < nchn) cur = chn[half * nt];
        if (half + 2 < nchn) nxt = chn[(half + 2) * nt];
    }
    for (int sl = threadIdx.x; sl < ns; sl += blockDim.x) {
        const int src = (sl < nt) ? c0 + sl : m.tile_halo[hb + sl - nt];
        for (int k = 0; k < 8; k += 2) {
            const double2 v = *reinterpret_cast<const double2*>(f.cs + plane_at(k, f.ld, src));
            *reinterpret_cast<double2*>(sq + ((k >> 1) * kSlots + sl) * 2) = v;
        }
        if constexpr (kSA) sq[8 * kSlots + sl] = sa.nu[src];
    }
    __syncthreads();
    const int c = c0 + cell;
    const LdsSlot<kSlots> self{sq, cell};
    double g[5][3], gs[3] = {0.0, 0.0, 0.0};
    for (int v = 0; v < 5; ++v) g[v][0] = g[v][1] = g[v][2] = 0.0;
    if (live) {
        double qc[5];
        for (int v = 0; v < 5; ++v) qc[v] = self[kRQ + v];
        const double x0 = self[kRC], y0 = self[kRC + 1], z0 = self[kRC + 2];
        const double sc0 = kSA ? sq[8 * kSlots + cell] : 0.0;
        const int cnt = m.ccn_off[c + 1] - m.ccn_off[c];
        const CellRef ri{f.cs, f.ld, c};
        const double i0 = ri[kRInv], i1 = ri[kRInv + 1], i2 = ri[kRInv + 2], i4 = ri[kRInv + 3], i5 = ri[kRInv + 4],
                     i8 = ri[kRInv + 5];
        for (int q = half; 4 * q < cnt; q += 2) {
            const SlotChunk ch = cur;
            cur = nxt;
            if (q + 4 < nchn) nxt = chn[(q + 4) * nt];
            for (int e = 0; e < 4 && 4 * q + e < cnt; ++e) {
                const int slot = chunk_slot(ch, e);
                const LdsSlot<kSlots> rn{sq, slot};
                const double dx = rn[kRC] - x0, dy = rn[kRC + 1] - y0, dz = rn[kRC + 2] - z0;
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double wgt = 1.0 / dmax(r2, 1.0e-30);
                const double rx = wgt * dx, ry = wgt * dy, rz = wgt * dz;
                const double cx = i0 * rx + i1 * ry + i2 * rz;
                const double cy = i1 * rx + i4 * ry + i5 * rz;
                const double cz = i2 * rx + i5 * ry + i8 * rz;
                for (int v = 0; v < 5; ++v) {
                    const double d = rn[kRQ + v] - qc[v];
                    g[v][0] += cx * d;
                    g[v][1] += cy * d;
                    g[v][2] += cz * d;
                }
                if constexpr (kSA) {
                    const double d = sq[8 * kSlots + slot] - sc0;
                    gs[0] += cx * d;
                    gs[1] += cy * d;
                    gs[2] += cz * d;
                }
            }
        }
    }
    // the second lane's sums added to the first lane's (every lane takes part)
    for (int v = 0; v < 5; ++v)
        for (int d = 0; d < 3; ++d) g[v][d] += __shfl_xor(g[v][d], 1, 64);
    if constexpr (kSA)
        for (int d = 0; d < 3; ++d) gs[d] += __shfl_xor(gs[d], 1, 64);
    if (!live || half != 0) return;
    if constexpr (kSA) {
        const long long nc = m.nc;
        sa.g[c] = gs[0];
        sa.g[nc + c] = gs[1];
        sa.g[2 * nc + c] = gs[2];
    }
    // the limiter's face neighbours and face vectors, in order, as the one-lane form reads them
    const int nchf = (m.ccf_base[t + 1] - m.ccf_base[t]) / (4 * nt);
    const SlotChunk* chf = reinterpret_cast<const SlotChunk*>(m.ccf_slot + m.ccf_base[t]) + cell;
    SlotChunk curf = 0, nxtf = 0;
    if (nchf > 0) curf = chf[0];
    if (nchf > 1) nxtf = chf[nt];
    const auto nbf = [&](int, int k) {
        if ((k & 3) == 0 && k > 0) {
            curf = nxtf;
            if ((k >> 2) + 1 < nchf) nxtf = chf[((k >> 2) + 1) * nt];
        }
        return LdsSlot<kSlots>{sq, chunk_slot(curf, k)};
    };
    const long long vb = m.face_vec_base[t];
    const long long vw = (m.face_vec_base[t + 1] - vb) / (3LL * nt);
    const double* fvp = m.face_vec + vb + cell;
    const auto fvec = [&](int, int k, double, double, double, double& ex, double& ey, double& ez) {
        ex = fvp[k * static_cast<long long>(nt)];
        ey = fvp[(vw + k) * nt];
        ez = fvp[(2 * vw + k) * nt];
    };
    store_and_limit(m, f, nm, c, self, g, nbf, fvec);
}

__device__ __forceinline__ void grad_dirichlet(double* g, const double* gi, double qin, double qfix, double nx,
                                               double ny, double nz, double rinv) {
    const double ddn = (qfix - qin) * rinv;
    const double gn = gi[0] * nx + gi[1] * ny + gi[2] * nz;
    const double fac = 2.0 * (ddn - gn);
    g[0] = gi[0] + fac * nx;
    g[1] = gi[1] + fac * ny;
    g[2] = gi[2] + fac * nz;
}

__device__ __forceinline__ void grad_neumann(double* g, const double* gi, double fixed, double nx, double ny,
                                             double nz) {
    const double gn = gi[0] * nx + gi[1] * ny + gi[2] * nz;
    const double fac = 2.0 * (fixed - gn);
    g[0] = gi[0] + fac * nx;
    g[1] = gi[1] + fac * ny;
    g[2] = gi[2] + fac * nz;
}

__device__ __forceinline__ void grad_slip(double* g, const double* gi, double qin, double qgh, double nx, double ny,
                                          double nz, double rinv) {
    const double gn = gi[0] * nx + gi[1] * ny + gi[2] * nz;
    const double ddn = 0.5 * (qgh - qin) * rinv;
    const double fac = 2.0 * (ddn - gn);
    g[0] = gi[0] + fac * nx;
    g[1] = gi[1] + fac * ny;
    g[2] = gi[2] + fac * nz;
}

// Ghost gradients (bc/physical/*.hpp gradient kernels, boundary_gradients.hpp)
__global__ void __launch_bounds__(kBlock) k_ghost_gradients(MeshArgs m, FieldArgs f, NumArgs nm, BcTable bt) {
    const int bf = blockIdx.x * blockDim.x + threadIdx.x;
    const int nbf = m.nf - m.ninner;
    if (bf >= nbf) return;
    const int fc = m.ninner + bf;
    const FaceRef fr = frec(m, fc);
    const CellRef ri = crec(f, face_owner(m, fc));
    const CellRefW rg = crecw(f, m.nc + bf);
    const BcDev bp = bt.bc[m.bface_bc[bf]];
    const double nx = fr[kFN], ny = fr[kFN + 1], nz = fr[kFN + 2];
    const double rinv = rcfn_inverse(fr, ri);
    double gi[5][3], go[5][3], qi[5];
    for (int v = 0; v < 5; ++v) {
        qi[v] = ri[kRQ + v];
        for (int d = 0; d < 3; ++d) go[v][d] = gi[v][d] = ri[kRG + 3 * v + d];
    }
    switch (bp.type) {
        case kSupIn: {
            const double qf[5] = {bp.p, bp.u, bp.v, bp.w, bp.T};
            for (int v = 0; v < 5; ++v) grad_dirichlet(go[v], gi[v], qi[v], qf[v], nx, ny, nz, rinv);
            break;
        }
        case kSlip:
        case kSym:
            grad_neumann(go[0], gi[0], 0.0, nx, ny, nz);
            for (int v = 1; v < 4; ++v) grad_slip(go[v], gi[v], qi[v], rg[kRQ + v], nx, ny, nz, rinv);
            grad_neumann(go[4], gi[4], 0.0, nx, ny, nz);
            break;
        case kFar: {
            double qb[5];
            bool so = false;
            farfield_state(bp, nm.gas.gamma, nm.gas.R, qi[0], qi[4], qi[1], qi[2], qi[3], nx, ny, nz, qb, so);
            if (!so)
                for (int v = 0; v < 5; ++v) grad_dirichlet(go[v], gi[v], qi[v], qb[v], nx, ny, nz, rinv);
            break;
        }
        case kNoSlip:
        case kSubIn: {
            const double qw[4] = {bp.u, bp.v, bp.w, bp.T};
            for (int v = 1; v < 5; ++v) grad_dirichlet(go[v], gi[v], qi[v], qw[v - 1], nx, ny, nz, rinv);
            break;
        }
        case kNoSlipHF: {
            const double qw[3] = {bp.u, bp.v, bp.w};
            for (int v = 1; v < 4; ++v) grad_dirichlet(go[v], gi[v], qi[v], qw[v - 1], nx, ny, nz, rinv);
            grad_neumann(go[4], gi[4], bp.tmp_grad, nx, ny, nz);
            break;
        }
        case kSubOut:
            grad_dirichlet(go[0], gi[0], qi[0], bp.p, nx, ny, nz, rinv);
            break;
        default:  // supersonic outlet: extrapolation
            break;
    }
    for (int v = 0; v < 5; ++v)
        for (int d = 0; d < 3; ++d) rg[kRG + 3 * v + d] = go[v][d];
}

// Inviscid + viscous flux of one face times its area, and its spectral-radius
// contribution (residual_kernel.hpp:183-240 / 245-300 for one face, with the
// MUSCL states of muscl.hpp:108-157).
// fr views the face's planes, (c0, c1) = (owner, neighbour or ghost).
// r0 / r1 view the owner's and the other cell's planes (in HBM or staged in
// LDS), mutL / mutR are their eddy viscosities (kEddy = false: none, both 0).
// The owner -> neighbour centroid vector and its inverse length are formed
// here with the operations of face_distances (aux_geometry.cpp:11-51,120-156),
// so they equal the mesh-order face planes bit for bit.
template <bool kEddy, class FR, class R0, class R1, class ML, class MR>
__device__ __forceinline__ void face_flux(const NumArgs& nm, const FR& fr, bool inner, const R0& r0, const R1& r1,
                                          const ML& mut0, const MR& mut1, double F[5], double& lam_f) {
    const double nx = fr[kFN], ny = fr[kFN + 1], nz = fr[kFN + 2], A = fr[kFA];
    double qL[5], qR[5];
    if (inner && nm.muscl) {
        const double fx = fr[kFC], fy = fr[kFC + 1], fz = fr[kFC + 2];
        const double d0x = fx - r0[kRC], d0y = fy - r0[kRC + 1], d0z = fz - r0[kRC + 2];
        const double d1x = fx - r1[kRC], d1y = fy - r1[kRC + 1], d1z = fz - r1[kRC + 2];
        for (int v = 0; v < 5; ++v) {
            const double dot0 = r0[kRG + 3 * v] * d0x + r0[kRG + 3 * v + 1] * d0y + r0[kRG + 3 * v + 2] * d0z;
            qL[v] = r0[kRQ + v] + r0[kRPhi + v] * dot0;
            const double dot1 = r1[kRG + 3 * v] * d1x + r1[kRG + 3 * v + 1] * d1y + r1[kRG + 3 * v + 2] * d1z;
            qR[v] = r1[kRQ + v] + r1[kRPhi + v] * dot1;
        }
    } else {
        for (int v = 0; v < 5; ++v) {
            qL[v] = r0[kRQ + v];
            qR[v] = r1[kRQ + v];
        }
    }
    double UL[5], UR[5];
    prim_to_cons(nm.gas, qL, UL);
    prim_to_cons(nm.gas, qR, UR);
    double smax = 0.0;
    hllc_flux(nm.gas, UL, UR, nx, ny, nz, A, F, smax);
    lam_f = A * smax;
    if (nm.viscous) {
        double qa[5], qb[5], ga[12], gb[12];
        for (int v = 0; v < 5; ++v) {
            qa[v] = r0[kRQ + v];
            qb[v] = r1[kRQ + v];
        }
        for (int k = 0; k < 12; ++k) {
            ga[k] = r0[kRG + 3 + k];
            gb[k] = r1[kRG + 3 + k];
        }
        double dvx, dvy, dvz, inv_d;
        if (inner) {
            const double ex = r1[kRC] - r0[kRC], ey = r1[kRC + 1] - r0[kRC + 1], ez = r1[kRC + 2] - r0[kRC + 2];
            dvx = ex;
            dvy = ey;
            dvz = ez;
            inv_d = 1.0 / sqrt(ex * ex + ey * ey + ez * ez);
        } else {
            const double ex = fr[kFC] - r0[kRC], ey = fr[kFC + 1] - r0[kRC + 1], ez = fr[kFC + 2] - r0[kRC + 2];
            dvx = 2.0 * ex;
            dvy = 2.0 * ey;
            dvz = 2.0 * ez;
            inv_d = 1.0 / sqrt(4.0 * ex * ex + 4.0 * ey * ey + 4.0 * ez * ez);
        }
        double Fv[5];
        double lv = 0.0;
        const double mutL = kEddy ? mut0() : 0.0, mutR = kEddy ? mut1() : 0.0;  // read where they are used
        viscous_flux<kEddy>(nm.gas, inv_d, dvx, dvy, dvz, qa, qb, ga, gb, mutL, mutR, nx, ny, nz, A, Fv, lv);
        for (int v = 1; v < 5; ++v) F[v] += Fv[v];
        lam_f += lv;
    }
}

// View of a tile cell's planes staged in LDS: planes 0 .. kFluxPlanes-1 in
// the pair layout of the cell planes with a stride of kTileCells slots.
constexpr int kFluxPlanes = 30;  // q, centroid, limiter, volume, gradients (plane 29 unused)
struct LdsRec {
    const double* s;
    int i;  // slot = device id - tile start
    __device__ __forceinline__ double operator[](int k) const { return s[((k >> 1) * kTileCells + i) * 2 + (k & 1)]; }
};

// One entry of the tile face-list stream (non-temporal loads); the centroid's
// z, which shares its 16-byte pair with the packed cell ids, is held by value.
struct TileFaceRef {
    const double* b;
    long long ld;
    int e;
    double fz;
    __device__ __forceinline__ double operator[](int k) const {
        if (k == kFC + 2) return fz;
        const double* base = b + static_cast<long long>(k >> 1) * 2 * ld;
        return __builtin_nontemporal_load(base + (2u * static_cast<unsigned>(e) + static_cast<unsigned>(k & 1)));
    }
};

// Tiled flux residual: each face touching the tile is evaluated once (face-
// parallel, one face per lane) and each cell then adds its faces' fluxes in
// ascending face index, with + for the owner and - for the neighbour
// (residual_kernel.hpp:175-306 formed as a gather; only faces on a tile
// boundary are evaluated by both adjacent tiles).  The tile's own cells are
// staged into LDS first (coalesced), so only the halo and ghost side of a
// face on the tile boundary is read from HBM; the face list is grouped by
// which cells lie in the tile, so a wave mostly takes one of the three load
// paths.  After the face phase the staging area is reused for the fluxes.
// With turbulence the kernel also records each face's mass flux F[0] (from
// the tile that owns the face's owner cell); eddy viscosity follows the
// cells (residual_kernel.hpp:229-230, 282-283).
template <bool kTurb>
__global__ void __launch_bounds__(kFluxThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) k_flux(MeshArgs m, FieldArgs f, NumArgs nm) {
    __shared__ double sc[kFluxPlanes * kTileCells];  // staged planes, then the fluxes sF[6][kTileFaces]
    __shared__ double smut[kTurb ? kTileCells : 1];
    __shared__ int scf[kTileCells + 1 + 2 * kTileFaces];  // the tile's cf_off (rebased), then its cf_slot entries
    static_assert(6 * kTileFaces <= kFluxPlanes * kTileCells, "flux area must fit the staging area");
    constexpr int kPerLane = kTileFaces / kFluxThreads;  // faces per lane
    const int t = xcd_block(blockIdx.x, gridDim.x);
    const int c0 = m.tile_off[t], nt = m.tile_off[t + 1] - c0;
    const int fb = m.tile_face_off[t], nfl = m.tile_face_off[t + 1] - fb;
    const int i = threadIdx.x;
    int cx[kPerLane], cy[kPerLane];
    double fz[kPerLane];
#pragma unroll
    for (int r = 0; r < kPerLane; ++r) {  // this lane's faces: cells (and index), issued before the staging loads
        const int j = i + r * kFluxThreads;
        cx[r] = cy[r] = 0;
        fz[r] = 0.0;
        if (j < nfl) {
            typedef double pair_t __attribute__((ext_vector_type(2)));
            const pair_t w = __builtin_nontemporal_load(
                reinterpret_cast<const pair_t*>(m.tf + plane_at(kTCells - 1, m.tld, fb + j)));
            const unsigned long long ids = static_cast<unsigned long long>(__double_as_longlong(w.y));
            fz[r] = w.x;
            cx[r] = static_cast<int>(static_cast<unsigned>(ids));
            cy[r] = static_cast<int>(static_cast<unsigned>(ids >> 32));
        }
    }
    // staging of the plane pairs the face fluxes read: q (pairs 0-2), centroid
    // and limiter (2-6) for MUSCL, gradients (7-14) for MUSCL or viscous terms
    {
        const int sl = i % kTileCells, g = i / kTileCells;
        constexpr int kGroups = kFluxThreads / kTileCells, kPairs = kFluxPlanes / 2;
        const int p_hi = (nm.muscl || nm.viscous) ? kPairs : 3;
        if (sl < nt) {
            for (int p = g; p < p_hi; p += kGroups) {
                if (p > 3 && p < 7 && !nm.muscl) continue;  // centroid (pairs 2-3) stays: viscous terms need it
                const double2 v = *reinterpret_cast<const double2*>(f.cs + plane_at(2 * p, f.ld, c0 + sl));
                *reinterpret_cast<double2*>(sc + (p * kTileCells + sl) * 2) = v;
            }
            if (kTurb && g == 0) smut[sl] = f.mut[c0 + sl];
        }
        // the cells' face slots for the gather: a contiguous range of cf_slot
        // (at most two entries per tile face); ne > 2 kTileFaces cannot occur
        const int e0 = m.cf_off[c0], ne = m.cf_off[c0 + nt] - e0;
        if (i <= nt) scf[i] = m.cf_off[c0 + i] - e0;
        for (int e = i; e < ne; e += kFluxThreads) scf[kTileCells + 1 + e] = m.cf_slot[e0 + e];
    }
    __syncthreads();
    double F[kPerLane][6];
#pragma unroll
    for (int r = 0; r < kPerLane; ++r) {
        const int j = i + r * kFluxThreads;
        for (int v = 0; v < 6; ++v) F[r][v] = 0.0;
        if (j >= nfl) break;  // later faces of the lane are out of range too
        const TileFaceRef fr{m.tf, m.tld, fb + j, fz[r]};
        const bool inner = cy[r] < m.nc;  // boundary faces: the other cell is a ghost
        const unsigned s0 = static_cast<unsigned>(cx[r] - c0), s1 = static_cast<unsigned>(cy[r] - c0);
        const bool in0 = s0 < static_cast<unsigned>(nt), in1 = s1 < static_cast<unsigned>(nt);
        double* Fr = F[r];
        // eddy viscosities from LDS (tile cells) or HBM (halo and ghosts)
        const int g0 = cx[r], g1 = cy[r];
        const auto lds0 = [&] { return smut[s0]; };
        const auto lds1 = [&] { return smut[s1]; };
        const auto hbm0 = [&] { return f.mut[g0]; };
        const auto hbm1 = [&] { return f.mut[g1]; };
        if (in0 && in1) {
            face_flux<kTurb>(nm, fr, inner, LdsRec{sc, static_cast<int>(s0)}, LdsRec{sc, static_cast<int>(s1)}, lds0,
                             lds1, Fr, Fr[5]);
        } else if (in0) {
            face_flux<kTurb>(nm, fr, inner, LdsRec{sc, static_cast<int>(s0)}, crec(f, g1), lds0, hbm1, Fr, Fr[5]);
        } else {
            face_flux<kTurb>(nm, fr, inner, crec(f, g0), LdsRec{sc, static_cast<int>(s1)}, hbm0, lds1, Fr, Fr[5]);
        }
        if (kTurb) f.mdot_tile[fb + j] = Fr[0];
        if (!inner)  // a boundary face lies in exactly one tile: single writer; ghost id = cells + boundary index
            for (int v = 0; v < 5; ++v) f.bflux[static_cast<long long>(cy[r] - m.nc) * 5 + v] = Fr[v];
    }
    __syncthreads();  // every read of the staged planes is done
    double* sF = sc;
#pragma unroll
    for (int r = 0; r < kPerLane; ++r) {
        const int j = i + r * kFluxThreads;
        if (j < nfl)
            for (int v = 0; v < 6; ++v) sF[v * kTileFaces + j] = F[r][v];
    }
    __syncthreads();
    if (i >= nt) return;
    const int c = c0 + i;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double lam = 0.0;
    for (int e = scf[i]; e < scf[i + 1]; ++e) {
        const int s = scf[kTileCells + 1 + e];
        if (s >= 0) {
            for (int v = 0; v < 5; ++v) acc[v] += sF[v * kTileFaces + s];
            lam += sF[5 * kTileFaces + s];
        } else {
            const int k = -s - 1;
            for (int v = 0; v < 5; ++v) acc[v] -= sF[v * kTileFaces + k];
            lam += sF[5 * kTileFaces + k];
        }
    }
    for (int v = 0; v < 5; ++v) __builtin_nontemporal_store(acc[v], f.res + static_cast<long long>(v) * m.nc + c);
    __builtin_nontemporal_store(lam, f.lam + c);
}

// Implicit local time step and diagonal shift (solver.hpp:256-287); unsteady
// = c0 / dt of dual time stepping (0 in steady mode)
__global__ void __launch_bounds__(kBlock) k_dt(int nc, const double* cs, long long ld, const double* lam, double cfl,
                                               double unsteady, double* dt, double* alpha) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const double vol = cs[plane_at(kRVol, ld, c)];
    const double inv_cfl = 1.0 / cfl;
    const double l = dmax(lam[c], 1.0e-14);
    dt[c] = cfl * vol / l;
    alpha[c] = (l * inv_cfl) + (vol * unsteady);
}

// Mesh-order face mass flux from the tile-order copy: entry of each face in
// the face list of the tile that holds its owner
__global__ void __launch_bounds__(kBlock) k_gather_mdot(int nf, const int* face_entry, const double* mdot_tile,
                                                        double* mdot) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    mdot[f] = mdot_tile[face_entry[f]];
}

// Output fields of the solution files (Solver::write_fields, solver.hpp:856-878):
// rho, u, v, w, p = (gamma - 1) rho e (state_conversions.hpp:12-25), Mach =
// |rho v| / (rho a) with a = sqrt(gamma p / rho); out = 6 planes of nc
__global__ void __launch_bounds__(kBlock) k_output_fields(int nc, Gas g, const double* U, double* out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double u[5];
    for (int v = 0; v < 5; ++v) u[v] = U[static_cast<long long>(v) * nc + c];
    const double r = u[0];
    const double inv_rho = 1.0 / r;
    const double q2 = (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) * (inv_rho * inv_rho);
    const double e = (u[4] * inv_rho) - 0.5 * q2;
    const double p = (g.gamma - 1.0) * r * e;
    const double a = sqrt(g.gamma * p / r);
    double* o = out + c;
    o[0] = r;
    o[static_cast<long long>(nc)] = u[1] / r;
    o[2LL * nc] = u[2] / r;
    o[3LL * nc] = u[3] / r;
    o[4LL * nc] = p;
    o[5LL * nc] = sqrt(u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / (r * a);
}

// Explicit local time step: alpha = dt / V = CFL / lambda (solver.hpp:278-285)
__global__ void __launch_bounds__(kBlock) k_dt_explicit(int nc, const double* cs, long long ld, const double* lam,
                                                        double cfl, double* dt, double* alpha) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const double l = dmax(lam[c], 1.0e-14);
    const double cfl_over_l = cfl / l;
    dt[c] = cfl_over_l * cs[plane_at(kRVol, ld, c)];
    alpha[c] = cfl_over_l;
}

// Forward Euler update U <- U - alpha R (update_ops.hpp:28-44)
__global__ void __launch_bounds__(kBlock) k_explicit_update(int nc, const double* alpha, const double* res, double* U) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const double a = alpha[c];
    for (int v = 0; v < 5; ++v) {
        const long long k = static_cast<long long>(v) * nc + c;
        U[k] = U[k] - a * res[k];
    }
}

// BDF pseudo-source of dual time stepping, R += V (c0 U + c1 U^n [+ c2 U^{n-1}])
// (solver.hpp:546-575); U^n, U^{n-1} as 5 planes of nc
__global__ void __launch_bounds__(kBlock) k_unsteady_source(int nc, const double* cs, long long ld, int bdf,
                                                            double inv_dt, const double* U, const double* Un,
                                                            const double* Unm1, double* res) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const double vol = cs[plane_at(kRVol, ld, c)];
    for (int v = 0; v < 5; ++v) {
        const long long k = static_cast<long long>(v) * nc + c;
        if (bdf == 2) {
            const double c0 = 1.5 * inv_dt, c1 = -2.0 * inv_dt, c2 = 0.5 * inv_dt;
            res[k] += vol * (c0 * U[k] + c1 * Un[k] + c2 * Unm1[k]);
        } else {
            const double c0 = 1.0 * inv_dt, c1 = -1.0 * inv_dt;
            res[k] += vol * (c0 * U[k] + c1 * Un[k]);
        }
    }
}

// Store one 5x5 block (values v, sign +1 or -1 applied as 0.0 + v / 0.0 - v,
// or as is) in the storage layout; bfloat16 rows go out as three 4-byte words.
template <class TS>
__device__ __forceinline__ void store_block(TS* dst, const double v[25], int sign) {
    for (int i = 0; i < 25; ++i) {
        const double x = sign > 0 ? 0.0 + v[i] : (sign < 0 ? 0.0 - v[i] : v[i]);
        dst[(i / 5) * BlockLayout<TS>::row + i % 5] = cvt<TS>(x);
    }
}
template <>
__device__ __forceinline__ void store_block<bf16>(bf16* dst, const double v[25], int sign) {
    std::uint32_t* w = reinterpret_cast<std::uint32_t*>(dst);
    for (int a = 0; a < 5; ++a) {
        std::uint32_t h[6];
        for (int q = 0; q < 5; ++q) {
            const double x = sign > 0 ? 0.0 + v[a * 5 + q] : (sign < 0 ? 0.0 - v[a * 5 + q] : v[a * 5 + q]);
            h[q] = cvt<bf16>(x).u;
        }
        h[5] = 0;
        for (int q = 0; q < 3; ++q) w[a * 3 + q] = h[2 * q] | (h[2 * q + 1] << 16);
    }
}

// Block of the storage type TS as storage words (8 bytes for FP64, 4 for FP32
// and BF16), the unit of the Jacobian kernel's coalesced block stores.
template <class TS>
struct BlockWords {
    typedef std::uint32_t word;
    static constexpr int n = BlockLayout<TS>::block * static_cast<int>(sizeof(TS)) / 4;
};
template <>
struct BlockWords<double> {
    typedef unsigned long long word;
    static constexpr int n = 25;
};

// The words of one block as store_block writes it (sign: +1 -> 0 + v,
// -1 -> 0 - v, 0 -> v); a BF16 block's padding word is zero.
template <class TS>
__device__ __forceinline__ void block_words(const double v[25], int sign, typename BlockWords<TS>::word* w);
template <>
__device__ __forceinline__ void block_words<double>(const double v[25], int sign, unsigned long long* w) {
    for (int k = 0; k < 25; ++k) {
        const double x = sign > 0 ? 0.0 + v[k] : (sign < 0 ? 0.0 - v[k] : v[k]);
        w[k] = static_cast<unsigned long long>(__double_as_longlong(x));
    }
}
template <>
__device__ __forceinline__ void block_words<float>(const double v[25], int sign, std::uint32_t* w) {
    for (int k = 0; k < 25; ++k) {
        const double x = sign > 0 ? 0.0 + v[k] : (sign < 0 ? 0.0 - v[k] : v[k]);
        w[k] = f32_bits(cvt<float>(x));
    }
}
template <>
__device__ __forceinline__ void block_words<bf16>(const double v[25], int sign, std::uint32_t* w) {
    for (int a = 0; a < 5; ++a) {
        std::uint32_t h[6];
        for (int q = 0; q < 5; ++q) {
            const double x = sign > 0 ? 0.0 + v[a * 5 + q] : (sign < 0 ? 0.0 - v[a * 5 + q] : v[a * 5 + q]);
            h[q] = cvt<bf16>(x).u;
        }
        h[5] = 0;
        for (int q = 0; q < 3; ++q) w[a * 3 + q] = h[2 * q] | (h[2 * q + 1] << 16);
    }
    w[15] = 0;
}

// J += Jv (plus) or J -= Jv for the dense viscous block Jv of
// viscous_jacobian_dense, touching only its 11 nonzero entries.  Adding the
// pattern's zeros would change no stored value: it can only turn a -0 entry
// into +0, and the diagonal sums start at +0 while every off-diagonal store
// forms 0 + v or 0 - v.
__device__ __forceinline__ void add_viscous_jacobian(const double m[9], bool plus, double J[25]) {
    const double d[11] = {m[0], m[3], m[1], m[3], m[2], m[3], m[4], m[5], m[6], m[7], m[8]};
    const int at[11] = {5, 6, 10, 12, 15, 18, 20, 21, 22, 23, 24};
    for (int k = 0; k < 11; ++k) J[at[k]] = plus ? J[at[k]] + d[k] : J[at[k]] - d[k];
}

// One face of a Jacobian row (jacobian_kernel.hpp:104-270 for one face): the
// row cell's state from LDS, the face from the tile face stream, the other
// cell from LDS (tile cell) or the cell planes (halo cell, ghost).  Adds the
// face's dF of the row's own side to dg (+ dFL as owner, - dFR as
// neighbour); at an interior face returns true with the other side's dF in J
// (dFR as owner, dFL as neighbour: the caller stores + J or - J).
template <bool kTurb>
__device__ __forceinline__ bool jacobian_face(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm,
                                              const double* sq, const double* smut, int i, int c0, int nt, int j,
                                              bool own, double dg[25], double J[25]) {
    typedef double pair_t __attribute__((ext_vector_type(2)));
    // the row cell's state, re-read from LDS per face (fewer live registers)
    double qc[5], Uc[5];
    for (int v = 0; v < 5; ++v) qc[v] = sq[v * kTileCells + i];
    prim_to_cons(nm.gas, qc, Uc);
    const pair_t n01 = *reinterpret_cast<const pair_t*>(m.tf + plane_at(kFN, m.tld, j));
    const pair_t n2a = *reinterpret_cast<const pair_t*>(m.tf + plane_at(kFN + 2, m.tld, j));
    const pair_t zid = *reinterpret_cast<const pair_t*>(m.tf + plane_at(kTCells - 1, m.tld, j));
    const double nx = n01.x, ny = n01.y, nz = n2a.x, A = n2a.y;
    const unsigned long long ids = static_cast<unsigned long long>(__double_as_longlong(zid.y));
    const int cx = static_cast<int>(static_cast<unsigned>(ids));
    const int cy = static_cast<int>(static_cast<unsigned>(ids >> 32));
    const bool inner = cy < m.nc;  // boundary faces: the other cell is a ghost
    const int oc = own ? cy : cx;
    const unsigned so = static_cast<unsigned>(oc - c0);
    const bool oin = so < static_cast<unsigned>(nt);
    double qo[5], xo[3];
    double mo = 0.0;
    if (oin) {
        for (int v = 0; v < 5; ++v) qo[v] = sq[v * kTileCells + so];
        for (int d = 0; d < 3; ++d) xo[d] = sq[(5 + d) * kTileCells + so];
        if (kTurb) mo = smut[so];
    } else {
        const CellRef r = crec(f, oc);
        for (int v = 0; v < 5; ++v) qo[v] = r[kRQ + v];
        for (int d = 0; d < 3; ++d) xo[d] = r[kRC + d];
        if (kTurb) mo = f.mut[oc];
    }
    double Uo[5];
    prim_to_cons(nm.gas, qo, Uo);
    double qL[5], qR[5], UL[5], UR[5];
    for (int v = 0; v < 5; ++v) {
        qL[v] = own ? qc[v] : qo[v];
        qR[v] = own ? qo[v] : qc[v];
        UL[v] = own ? Uc[v] : Uo[v];
        UR[v] = own ? Uo[v] : Uc[v];
    }
    const double smax = hllc_smax(nm.gas, UL, UR, nx, ny, nz);
    const double half_area = 0.5 * A;
    const double diss = 0.5 * smax * A;
    double vm[9];
    if (nm.viscous) {
        // owner -> neighbour centroid vector and its inverse length with the
        // operations of face_distances (aux_geometry.cpp:11-51, 120-156)
        const double xc0 = sq[5 * kTileCells + i], xc1 = sq[6 * kTileCells + i], xc2 = sq[7 * kTileCells + i];
        double inv_d;
        if (inner) {
            const double ax = own ? xc0 : xo[0], ay = own ? xc1 : xo[1], az = own ? xc2 : xo[2];
            const double bx = own ? xo[0] : xc0, by = own ? xo[1] : xc1, bz = own ? xo[2] : xc2;
            const double ex = bx - ax, ey = by - ay, ez = bz - az;
            inv_d = 1.0 / sqrt(ex * ex + ey * ey + ez * ez);
        } else {  // the row cell owns a boundary face
            const pair_t fxy = *reinterpret_cast<const pair_t*>(m.tf + plane_at(kFC, m.tld, j));
            const double ex = fxy.x - xc0, ey = fxy.y - xc1, ez = zid.x - xc2;
            inv_d = 1.0 / sqrt(4.0 * ex * ex + 4.0 * ey * ey + 4.0 * ez * ez);
        }
        // jacobian_kernel.hpp:156-157, 209-210, 252-254: the ghost side carries no eddy viscosity
        const double mc = kTurb ? smut[i] : 0.0;
        const double mutL = kTurb ? (own ? mc : mo) : 0.0;
        const double mutR = (kTurb && inner) ? (own ? mo : mc) : 0.0;
        viscous_jacobian_entries<kTurb>(nm.gas, inv_d, qL, qR, mutL, mutR, A, vm);
    }
    // own side, dFL = 0.5 A J(UL) + diss I (+ viscous) as owner, dFR = 0.5 A
    // J(UR) - diss I (- viscous) as neighbour
    euler_jacobian_scaled(nm.gas, Uc, nx, ny, nz, half_area, J);
    for (int k = 0; k < 5; ++k) J[k * 6] = own ? (J[k * 6] + diss) : (J[k * 6] - diss);
    if (nm.viscous) add_viscous_jacobian(vm, own, J);
    if (own)
        for (int k = 0; k < 25; ++k) dg[k] += J[k];
    else
        for (int k = 0; k < 25; ++k) dg[k] -= J[k];
    if (!inner) return false;
    // other side: dFR as owner, dFL as neighbour
    euler_jacobian_scaled(nm.gas, Uo, nx, ny, nz, half_area, J);
    for (int k = 0; k < 5; ++k) J[k * 6] = own ? (J[k * 6] - diss) : (J[k * 6] + diss);
    if (nm.viscous) add_viscous_jacobian(vm, !own, J);
    return true;
}

// Picard Jacobian on the flux tiles (jacobian_kernel.hpp:74-270): one
// workgroup per tile, one thread per block row.  The tile's cells' primitive
// states, centroids (and eddy viscosities) are staged in LDS with coalesced
// loads; each of a row's faces is read from the tile face-list stream (normal,
// area, centroid, packed cell ids: contiguous per tile), so only the halo or
// ghost side of a face on the tile boundary is gathered from the cell planes.
// Row c forms, per face in ascending face index, the face's dF of its own
// side (added to the diagonal: + dFL as owner, - dFR as neighbour) and, at an
// interior face, the other side's dF, the off-diagonal block at the position
// precomputed at set-up (cf_blk): + dFR as owner, - dFL as neighbour.
// The rows advance through their faces in lock step (round r = each row's
// r-th face, then a last round for the diagonal blocks); each round's blocks
// go through LDS, and each wave stores its 64 blocks word by word with
// consecutive lanes on consecutive words, so a store instruction writes a few
// contiguous runs instead of 64 scattered words.
// The arithmetic and its order are those of the reference's face loop, so the
// blocks are the reference's bit for bit; each block has a single writer.
// Waves per SIMD: three where the LDS leaves room for three workgroups per
// CU (FP32 blocks without eddy viscosity, BF16 blocks: 52 KB and 45 KB), at
// 168 registers; two otherwise (FP64 blocks: 78 KB; FP32 with eddy
// viscosity: 54 KB), where a lower register cap would only add spills.
template <class TS, bool kTurb>
__global__ void __launch_bounds__(kTileCells)
__attribute__((amdgpu_waves_per_eu((sizeof(TS) == 8 || (sizeof(TS) == 4 && kTurb)) ? 2 : 3)))
k_jacobian(MeshArgs m, FieldArgs f, NumArgs nm, BsrArgs b,
                                                         const double* alpha, TS* values) {
    typedef typename BlockWords<TS>::word word;
    constexpr int kW = BlockWords<TS>::n;
    __shared__ double sq[8 * kTileCells];  // q (0-4) and centroid (5-7) of the tile's cells, one plane each
    __shared__ double smut[kTurb ? kTileCells : 1];
    __shared__ int scf[kTileCells + 1 + 2 * kTileFaces];  // the tile's cf_off (rebased), then its cf_slot entries
    __shared__ word sblk[kTileCells * kW];  // one block per row for the round's store
    __shared__ int spos[kTileCells];        // its destination block, -1 for none
    __shared__ int srounds;
    const int t = xcd_block(blockIdx.x, gridDim.x);
    const int c0 = m.tile_off[t], nt = m.tile_off[t + 1] - c0;
    const int fb = m.tile_face_off[t];
    const int i = threadIdx.x;
    const int e0 = m.cf_off[c0];
    if (i == 0) srounds = 0;
    if (i < nt) {
        typedef double pair_t __attribute__((ext_vector_type(2)));
        for (int p = 0; p < 4; ++p) {
            const pair_t v = *reinterpret_cast<const pair_t*>(f.cs + plane_at(2 * p, f.ld, c0 + i));
            sq[(2 * p) * kTileCells + i] = v.x;
            sq[(2 * p + 1) * kTileCells + i] = v.y;
        }
        if (kTurb) smut[i] = f.mut[c0 + i];
    }
    {
        // nt + 1 offsets (up to kTileCells + 1, one more than the threads)
        const int ne = m.cf_off[c0 + nt] - e0;  // at most two entries per tile face
        for (int k = i; k <= nt; k += kTileCells) scf[k] = m.cf_off[c0 + k] - e0;
        for (int e = i; e < ne; e += kTileCells) scf[kTileCells + 1 + e] = m.cf_slot[e0 + e];
    }
    __syncthreads();
    const bool row = i < nt;
    const int nfc = row ? scf[i + 1] - scf[i] : 0;
    if (nfc > 0) atomicMax(&srounds, nfc);  // integer maximum: the tile's longest face list
    const int c = c0 + (row ? i : 0);
    double dg[25];
    for (int k = 0; k < 25; ++k) dg[k] = 0.0;
    __syncthreads();
    const int rounds = srounds;
    const int lane = i & 63, wbase = i & ~63;
    for (int rr = 0; rr <= rounds; ++rr) {
        int pos = -1;
        if (rr < nfc) {
            const int e = scf[i] + rr;
            const int sl = scf[kTileCells + 1 + e];
            const bool own = sl >= 0;
            double J[25];
            if (jacobian_face<kTurb>(m, f, nm, sq, smut, i, c0, nt, fb + (own ? sl : -sl - 1), own, dg, J)) {
                pos = b.cf_blk[e0 + e];
                block_words<TS>(J, own ? 1 : -1, sblk + i * kW);
            }
        } else if (rr == rounds && row) {  // last round: the diagonal block
            const double a = alpha[c];
            for (int k = 0; k < 5; ++k) dg[k * 6] += a;
            pos = b.diag[c];
            block_words<TS>(dg, 0, sblk + i * kW);
        }
        spos[i] = pos;
        __syncthreads();
        // this wave's 64 blocks, word by word: consecutive lanes on consecutive words
        word* out = reinterpret_cast<word*>(values);
#pragma unroll 1
        for (int q = 0; q < kW; ++q) {
            const int w = q * 64 + lane;
            const int blk = w / kW, o = w - blk * kW;
            const int dst = spos[wbase + blk];
            if (dst >= 0) out[static_cast<long long>(dst) * kW + o] = sblk[(wbase + blk) * kW + o];
        }
        __syncthreads();  // the next round reuses sblk / spos
    }
}

// Per-patch boundary integrals (residual_kernel.hpp:96-169 forms F[0] and
// F[4]; the momentum components give the force on the patch).  One workgroup
// per patch; thread t sums the patch's faces t, t + 256, ... in order, then a
// fixed tree adds the thread sums.
__global__ void __launch_bounds__(kBlock) k_boundary_integrals(int nbf, const int* bface_bc, const double* bflux,
                                                               double* out) {
    __shared__ double red[kBlock];
    const int p = blockIdx.x;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nbf; b += kBlock) {
        if (bface_bc[b] != p) continue;
        for (int v = 0; v < 5; ++v) acc[v] += bflux[static_cast<long long>(b) * 5 + v];
    }
    for (int v = 0; v < 5; ++v) {
        red[threadIdx.x] = acc[v];
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {
            if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[p * 5 + v] = red[0];
        __syncthreads();
    }
}

inline dim3 grid_for(long long n) { return dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)); }

}  // namespace

void launch_primitives(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, hipStream_t s) {
    if (m.nc > 0) hipLaunchKernelGGL(k_primitives, grid_for(m.nc), dim3(kBlock), 0, s, m, f, nm);
}

void launch_ghosts(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, const BcTable& bcs, hipStream_t s) {
    const int nbf = m.nf - m.ninner;
    if (nbf > 0) hipLaunchKernelGGL(k_ghosts, grid_for(nbf), dim3(kBlock), 0, s, m, f, nm, bcs);
}

bool launch_gradient_limiter(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, hipStream_t s,
                             const SaArgs* sa) {
    if (m.nc <= 0) return false;
    if (m.tile_halo_off != nullptr) {
        const dim3 g(static_cast<unsigned>(m.ngtiles)), b(static_cast<unsigned>(m.gblock));
        const bool big = m.gslots > kGradSlots;
        const dim3 b2(2u * static_cast<unsigned>(m.gblock));  // the wide form: two lanes per cell
        if (sa) {
            if (big)
                hipLaunchKernelGGL((k_gradient_limiter_pair<true>), g, b2, 0, s, m, f, nm, *sa);
            else
                hipLaunchKernelGGL((k_gradient_limiter_lds<true, kGradSlots>), g, b, 0, s, m, f, nm, *sa);
            return true;
        }
        if (big)
            hipLaunchKernelGGL((k_gradient_limiter_pair<false>), g, b2, 0, s, m, f, nm, SaArgs{});
        else
            hipLaunchKernelGGL((k_gradient_limiter_lds<false, kGradSlots>), g, b, 0, s, m, f, nm, SaArgs{});
        return false;
    }
    hipLaunchKernelGGL(k_gradient_limiter, grid_for(m.nc), dim3(kBlock), 0, s, m, f, nm);
    return false;
}

void launch_lsq_setup(const MeshArgs& m, const FieldArgs& f, hipStream_t s) {
    if (m.nc > 0) hipLaunchKernelGGL(k_lsq_setup, grid_for(m.nc), dim3(kBlock), 0, s, m, f);
}

void launch_ghost_gradients(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, const BcTable& bcs,
                            hipStream_t s) {
    const int nbf = m.nf - m.ninner;
    if (nbf > 0) hipLaunchKernelGGL(k_ghost_gradients, grid_for(nbf), dim3(kBlock), 0, s, m, f, nm, bcs);
}

void launch_flux(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, hipStream_t s) {
    if (m.ntiles <= 0) return;
    if (f.mut)
        hipLaunchKernelGGL(k_flux<true>, dim3(static_cast<unsigned>(m.ntiles)), dim3(kFluxThreads), 0, s, m, f, nm);
    else
        hipLaunchKernelGGL(k_flux<false>, dim3(static_cast<unsigned>(m.ntiles)), dim3(kFluxThreads), 0, s, m, f, nm);
}

void launch_boundary_integrals(const MeshArgs& m, const FieldArgs& f, int n_patches, double* out, hipStream_t s) {
    if (n_patches > 0)
        hipLaunchKernelGGL(k_boundary_integrals, dim3(static_cast<unsigned>(n_patches)), dim3(kBlock), 0, s,
                           m.nf - m.ninner, m.bface_bc, f.bflux, out);
}

void launch_output_fields(int nc, const Gas& g, const double* U, double* out, hipStream_t s) {
    if (nc > 0) hipLaunchKernelGGL(k_output_fields, grid_for(nc), dim3(kBlock), 0, s, nc, g, U, out);
}

void launch_gather_mdot(int nf, const int* face_entry, const double* mdot_tile, double* mdot, hipStream_t s) {
    if (nf > 0) hipLaunchKernelGGL(k_gather_mdot, grid_for(nf), dim3(kBlock), 0, s, nf, face_entry, mdot_tile, mdot);
}

void launch_dt(const MeshArgs& m, const FieldArgs& f, const double* lam, double cfl, double unsteady, double* dt,
               double* alpha, hipStream_t s) {
    if (m.nc > 0)
        hipLaunchKernelGGL(k_dt, grid_for(m.nc), dim3(kBlock), 0, s, m.nc, f.cs, f.ld, lam, cfl, unsteady, dt, alpha);
}

void launch_dt_explicit(const MeshArgs& m, const FieldArgs& f, const double* lam, double cfl, double* dt,
                        double* alpha, hipStream_t s) {
    if (m.nc > 0)
        hipLaunchKernelGGL(k_dt_explicit, grid_for(m.nc), dim3(kBlock), 0, s, m.nc, f.cs, f.ld, lam, cfl, dt, alpha);
}

void launch_explicit_update(const MeshArgs& m, const FieldArgs& f, const double* alpha, double* U, hipStream_t s) {
    if (m.nc > 0) hipLaunchKernelGGL(k_explicit_update, grid_for(m.nc), dim3(kBlock), 0, s, m.nc, alpha, f.res, U);
}

void launch_unsteady_source(const MeshArgs& m, const FieldArgs& f, int bdf, double inv_dt, const double* Un,
                            const double* Unm1, hipStream_t s) {
    if (m.nc > 0)
        hipLaunchKernelGGL(k_unsteady_source, grid_for(m.nc), dim3(kBlock), 0, s, m.nc, f.cs, f.ld, bdf, inv_dt, f.U,
                           Un, Unm1, f.res);
}

void launch_jacobian(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, const BsrArgs& b,
                     const double* alpha, void* values, int storage, hipStream_t s) {
    if (m.nc <= 0) return;
    const dim3 grid(static_cast<unsigned>(m.ntiles)), block(kTileCells);
    const bool turb = f.mut != nullptr;
    if (storage == kStoreF32) {
        float* v = static_cast<float*>(values);
        if (turb) hipLaunchKernelGGL((k_jacobian<float, true>), grid, block, 0, s, m, f, nm, b, alpha, v);
        else hipLaunchKernelGGL((k_jacobian<float, false>), grid, block, 0, s, m, f, nm, b, alpha, v);
    } else if (storage == kStoreBF16) {
        bf16* v = static_cast<bf16*>(values);
        if (turb) hipLaunchKernelGGL((k_jacobian<bf16, true>), grid, block, 0, s, m, f, nm, b, alpha, v);
        else hipLaunchKernelGGL((k_jacobian<bf16, false>), grid, block, 0, s, m, f, nm, b, alpha, v);
    } else {
        double* v = static_cast<double*>(values);
        if (turb) hipLaunchKernelGGL((k_jacobian<double, true>), grid, block, 0, s, m, f, nm, b, alpha, v);
        else hipLaunchKernelGGL((k_jacobian<double, false>), grid, block, 0, s, m, f, nm, b, alpha, v);
    }
}

}  // namespace cfdg
