// Residual, time-step and Jacobian kernels (gfx950).
//
// Design: every kernel is cell-centric (one thread per cell, or per boundary
// face for the ghost stages).  Face contributions are gathered over the
// cell's face list in ascending face index -- the order in which the
// reference's face sweeps scatter into a cell -- so each cell sum is formed
// in the same order, without atomics, and is reproducible run to run.  The
// price is that an interior face's flux is evaluated by both adjacent cells.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace cfdg {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ const double* plane(const double* a, int ld, int v) { return a + static_cast<long long>(v) * ld; }
__device__ __forceinline__ double* plane(double* a, int ld, int v) { return a + static_cast<long long>(v) * ld; }

// ---------------------------------------------------------------------------
// U -> q for the cells (solver.hpp:185-188)
__global__ void __launch_bounds__(kBlock) k_primitives(MeshArgs m, FieldArgs f, NumArgs nm) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nc) return;
    double U[5], q[5];
    for (int v = 0; v < 5; ++v) U[v] = plane(f.U, m.ld, v)[c];
    cons_to_prim(nm.gas, U, q);
    for (int v = 0; v < 5; ++v) plane(f.q, m.ld, v)[c] = q[v];
}

__device__ __forceinline__ double rcfn_inverse(const MeshArgs& m, int f, int in) {
    // boundary_traversal.hpp:98-108
    const double rx = m.fcx[f] - m.ccx[in], ry = m.fcy[f] - m.ccy[in], rz = m.fcz[f] - m.ccz[in];
    const double r = rx * m.fnx[f] + ry * m.fny[f] + rz * m.fnz[f];
    return 1.0 / dmax(r, 1.0e-14);
}

// Ghost primitive states, one thread per boundary face (bc/physical/*.hpp value kernels)
__global__ void __launch_bounds__(kBlock) k_ghosts(MeshArgs m, FieldArgs f, NumArgs nm, BcTable bt) {
    const int bf = blockIdx.x * blockDim.x + threadIdx.x;
    const int nbf = m.nf - m.ninner;
    if (bf >= nbf) return;
    const int fc = m.ninner + bf;
    const int in = m.owner[fc];
    const int gh = m.nc + bf;
    const BcDev bp = bt.bc[m.bface_bc[bf]];
    double qi[5], qg[5];
    for (int v = 0; v < 5; ++v) qi[v] = plane(f.q, m.ld, v)[in];
    const double nx = m.fnx[fc], ny = m.fny[fc], nz = m.fnz[fc];
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
        case kSubIn: {
            qg[0] = qi[0];
            qg[1] = 2.0 * bp.u - qi[1];
            qg[2] = 2.0 * bp.v - qi[2];
            qg[3] = 2.0 * bp.w - qi[3];
            qg[4] = 2.0 * bp.T - qi[4];
            qg[4] = dmax(qg[4], 1.0);
            break;
        }
        case kNoSlipHF: {
            const double rinv = rcfn_inverse(m, fc, in);
            qg[0] = qi[0];
            qg[1] = 2.0 * bp.u - qi[1];
            qg[2] = 2.0 * bp.v - qi[2];
            qg[3] = 2.0 * bp.w - qi[3];
            const double r = 1.0 / rinv;
            qg[4] = qi[4] + 2.0 * bp.tmp_grad * r;
            qg[4] = dmax(qg[4], 1.0);
            break;
        }
        case kSubOut: {
            qg[0] = 2.0 * bp.p - qi[0];
            qg[0] = dmax(qg[0], 1.0);
            for (int v = 1; v < 5; ++v) qg[v] = qi[v];
            break;
        }
        default:  // kSupOut: extrapolation
            for (int v = 0; v < 5; ++v) qg[v] = qi[v];
            break;
    }
    for (int v = 0; v < 5; ++v) plane(f.q, m.ld, v)[gh] = qg[v];
}

// Cell gradients (Green-Gauss face/cell or node-stencil least squares) fused
// with the MUSCL limiter of the same cell (muscl.hpp:26-104): the limiter only
// needs the cell's own gradient, which stays in registers.
__global__ void __launch_bounds__(kBlock) k_gradient_limiter(MeshArgs m, FieldArgs f, NumArgs nm) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nc) return;
    double qc[5];
    for (int v = 0; v < 5; ++v) qc[v] = plane(f.q, m.ld, v)[c];
    double g[5][3];
    for (int v = 0; v < 5; ++v) g[v][0] = g[v][1] = g[v][2] = 0.0;

    if (nm.gradient == 2) {
        // gradient_operator.cpp:549-599
        const double* wx = m.lsq;
        const double* wy = m.lsq + m.n_edges;
        const double* wz = m.lsq + 2 * m.n_edges;
        for (int j = m.ccn_off[c]; j < m.ccn_off[c + 1]; ++j) {
            const int nb = m.ccn[j];
            const double cx = wx[j], cy = wy[j], cz = wz[j];
            for (int v = 0; v < 5; ++v) {
                const double d = plane(f.q, m.ld, v)[nb] - qc[v];
                g[v][0] += cx * d;
                g[v][1] += cy * d;
                g[v][2] += cz * d;
            }
        }
    } else {
        // Green-Gauss: face scatter of gradient_operator.cpp:197-278 gathered per cell
        // (the cell-based variant, :61-130, forms the same sums)
        for (int e = m.cf_off[c]; e < m.cf_off[c + 1]; ++e) {
            const int fc = m.cf[e];
            const int o = m.owner[fc];
            const int nb = (fc < m.ninner) ? m.neigh[fc] : m.nc + (fc - m.ninner);
            const double a05 = 0.5 * m.farea[fc];
            const double hx = a05 * m.fnx[fc], hy = a05 * m.fny[fc], hz = a05 * m.fnz[fc];
            const bool own = (o == c);
            for (int v = 0; v < 5; ++v) {
                const double s = plane(f.q, m.ld, v)[o] + plane(f.q, m.ld, v)[nb];
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
        const double iv = 1.0 / m.vol[c];
        for (int v = 0; v < 5; ++v) {
            g[v][0] *= iv;
            g[v][1] *= iv;
            g[v][2] *= iv;
        }
    }
    for (int v = 0; v < 5; ++v)
        for (int d = 0; d < 3; ++d) plane(f.grad, m.ld, 3 * v + d)[c] = g[v][d];

    if (!nm.muscl) return;
    double qmax[5], qmin[5];
    for (int v = 0; v < 5; ++v) qmax[v] = qmin[v] = qc[v];
    for (int j = m.ccf_off[c]; j < m.ccf_off[c + 1]; ++j) {
        const int nb = m.ccf[j];
        for (int v = 0; v < 5; ++v) {
            const double qn = plane(f.q, m.ld, v)[nb];
            qmax[v] = dmax(qmax[v], qn);
            qmin[v] = dmin(qmin[v], qn);
        }
    }
    double ph[5] = {1.0, 1.0, 1.0, 1.0, 1.0};
    const double eps2 = nm.k3 * m.vol[c];
    const double cx = m.ccx[c], cy = m.ccy[c], cz = m.ccz[c];
    for (int e = m.cf_off[c]; e < m.cf_off[c + 1]; ++e) {
        const int fc = m.cf[e];
        const double ex = m.fcx[fc] - cx, ey = m.fcy[fc] - cy, ez = m.fcz[fc] - cz;
        for (int v = 0; v < 5; ++v) {
            const double df = g[v][0] * ex + g[v][1] * ey + g[v][2] * ez;
            const double dnb = (df > 0.0) ? (qmax[v] - qc[v]) : (qmin[v] - qc[v]);
            ph[v] = dmin(ph[v], limiter_phi(nm.limiter, dnb, df, eps2));
        }
    }
    for (int v = 0; v < 5; ++v) plane(f.phi, m.ld, v)[c] = ph[v];
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

// Ghost gradients, one thread per boundary face (bc/physical/*.hpp gradient kernels,
// numerics/boundary/boundary_gradients.hpp)
__global__ void __launch_bounds__(kBlock) k_ghost_gradients(MeshArgs m, FieldArgs f, NumArgs nm, BcTable bt) {
    const int bf = blockIdx.x * blockDim.x + threadIdx.x;
    const int nbf = m.nf - m.ninner;
    if (bf >= nbf) return;
    const int fc = m.ninner + bf;
    const int in = m.owner[fc];
    const int gh = m.nc + bf;
    const BcDev bp = bt.bc[m.bface_bc[bf]];
    const double nx = m.fnx[fc], ny = m.fny[fc], nz = m.fnz[fc];
    const double rinv = rcfn_inverse(m, fc, in);
    double gi[5][3], go[5][3], qi[5];
    for (int v = 0; v < 5; ++v) {
        qi[v] = plane(f.q, m.ld, v)[in];
        for (int d = 0; d < 3; ++d) go[v][d] = gi[v][d] = plane(f.grad, m.ld, 3 * v + d)[in];
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
            for (int v = 1; v < 4; ++v) grad_slip(go[v], gi[v], qi[v], plane(f.q, m.ld, v)[gh], nx, ny, nz, rinv);
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
        default:  // kSupOut: extrapolation
            break;
    }
    for (int v = 0; v < 5; ++v)
        for (int d = 0; d < 3; ++d) plane(f.grad, m.ld, 3 * v + d)[gh] = go[v][d];
}

// Flux residual: R_c = sum over the cell's faces of +-(F_inv + F_visc) A and
// lambda_c = sum of A s_max + lambda_visc (residual_kernel.hpp:175-306), with
// MUSCL face states (muscl.hpp:108-157).
__global__ void __launch_bounds__(kBlock) k_flux(MeshArgs m, FieldArgs f, NumArgs nm) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nc) return;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double lam = 0.0;
    for (int e = m.cf_off[c]; e < m.cf_off[c + 1]; ++e) {
        const int fc = m.cf[e];
        const bool inner = fc < m.ninner;
        const int c0 = m.owner[fc];
        const int c1 = inner ? m.neigh[fc] : m.nc + (fc - m.ninner);
        const double nx = m.fnx[fc], ny = m.fny[fc], nz = m.fnz[fc], A = m.farea[fc];
        double qa[5], qb[5];
        for (int v = 0; v < 5; ++v) {
            qa[v] = plane(f.q, m.ld, v)[c0];
            qb[v] = plane(f.q, m.ld, v)[c1];
        }
        double qL[5], qR[5];
        if (inner && nm.muscl) {
            const double fx = m.fcx[fc], fy = m.fcy[fc], fz = m.fcz[fc];
            const double d0x = fx - m.ccx[c0], d0y = fy - m.ccy[c0], d0z = fz - m.ccz[c0];
            const double d1x = fx - m.ccx[c1], d1y = fy - m.ccy[c1], d1z = fz - m.ccz[c1];
            for (int v = 0; v < 5; ++v) {
                const double dot0 = plane(f.grad, m.ld, 3 * v)[c0] * d0x + plane(f.grad, m.ld, 3 * v + 1)[c0] * d0y +
                                    plane(f.grad, m.ld, 3 * v + 2)[c0] * d0z;
                qL[v] = qa[v] + plane(f.phi, m.ld, v)[c0] * dot0;
                const double dot1 = plane(f.grad, m.ld, 3 * v)[c1] * d1x + plane(f.grad, m.ld, 3 * v + 1)[c1] * d1y +
                                    plane(f.grad, m.ld, 3 * v + 2)[c1] * d1z;
                qR[v] = qb[v] + plane(f.phi, m.ld, v)[c1] * dot1;
            }
        } else {
            for (int v = 0; v < 5; ++v) {
                qL[v] = qa[v];
                qR[v] = qb[v];
            }
        }
        double UL[5], UR[5], F[5];
        prim_to_cons(nm.gas, qL, UL);
        prim_to_cons(nm.gas, qR, UR);
        double smax = 0.0;
        hllc_flux(nm.gas, UL, UR, nx, ny, nz, A, F, smax);
        double lam_f = A * smax;
        if (nm.viscous) {
            double ga[12], gb[12];
            for (int k = 0; k < 12; ++k) {
                ga[k] = plane(f.grad, m.ld, 3 + k)[c0];
                gb[k] = plane(f.grad, m.ld, 3 + k)[c1];
            }
            double Fv[5];
            double lv = 0.0;
            viscous_flux(nm.gas, m.dinv[fc], m.dvx[fc], m.dvy[fc], m.dvz[fc], qa, qb, ga, gb, 0.0, 0.0, nx, ny, nz, A,
                         Fv, lv);
            for (int v = 1; v < 5; ++v) F[v] += Fv[v];
            lam_f += lv;
        }
        if (c0 == c) {
            for (int v = 0; v < 5; ++v) acc[v] += F[v];
        } else {
            for (int v = 0; v < 5; ++v) acc[v] -= F[v];
        }
        lam += lam_f;
    }
    for (int v = 0; v < 5; ++v) plane(f.res, m.ld, v)[c] = acc[v];
    f.lam[c] = lam;
}

// Implicit local time step and diagonal shift (solver.hpp:256-287)
__global__ void __launch_bounds__(kBlock) k_dt(int nc, const double* vol, const double* lam, double cfl, double* dt,
                                               double* alpha) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const double inv_cfl = 1.0 / cfl;
    const double l = dmax(lam[c], 1.0e-14);
    dt[c] = cfl * vol[c] / l;
    alpha[c] = (l * inv_cfl) + (vol[c] * 0.0);
}

__device__ __forceinline__ int find_col(const BsrArgs& b, int row, int col) {
    int k = b.row_ptr[row];
    const int e = b.row_ptr[row + 1];
    while (k < e && b.cols[k] != col) ++k;
    return k;  // the pattern holds every face neighbour
}

// First-order Picard Jacobian, one thread per block row (jacobian_kernel.hpp:74-270):
// every block of row c is produced by the thread of row c, the diagonal block
// accumulated over the cell's faces in ascending face order.
__global__ void __launch_bounds__(kBlock) k_jacobian(MeshArgs m, FieldArgs f, NumArgs nm, BsrArgs b,
                                                     const double* alpha, double* values) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nc) return;
    double dg[25];
    for (int i = 0; i < 25; ++i) dg[i] = 0.0;
    for (int e = m.cf_off[c]; e < m.cf_off[c + 1]; ++e) {
        const int fc = m.cf[e];
        const bool inner = fc < m.ninner;
        const int c0 = m.owner[fc];
        const int c1 = inner ? m.neigh[fc] : m.nc + (fc - m.ninner);
        const double nx = m.fnx[fc], ny = m.fny[fc], nz = m.fnz[fc], A = m.farea[fc];
        double qL[5], qR[5], UL[5], UR[5];
        for (int v = 0; v < 5; ++v) {
            qL[v] = plane(f.q, m.ld, v)[c0];
            qR[v] = plane(f.q, m.ld, v)[c1];
        }
        prim_to_cons(nm.gas, qL, UL);
        prim_to_cons(nm.gas, qR, UR);
        const double smax = hllc_smax(nm.gas, UL, UR, nx, ny, nz);
        const double half_area = 0.5 * A;
        const double diss = 0.5 * smax * A;
        double vm[9];
        if (nm.viscous) viscous_jacobian_entries(nm.gas, m.dinv[fc], qL, qR, 0.0, 0.0, A, vm);
        double J[25], Jv[25];
        const bool own = (c0 == c);
        // Left (owner) side: dFL = 0.5 A J(UL) + diss I (+ viscous); right side:
        // dFR = 0.5 A J(UR) - diss I (- viscous).
        // Row c needs: diag += dFL (owner) or -= dFR (neighbour); off-diagonal =
        // dFR (owner) or -dFL (neighbour).
        for (int pass = 0; pass < (inner ? 2 : 1); ++pass) {
            const bool left = (pass == 0) ? own : !own;  // pass 0: diagonal term, pass 1: off-diagonal term
            euler_jacobian_scaled(nm.gas, left ? UL : UR, nx, ny, nz, half_area, J);
            for (int i = 0; i < 5; ++i) J[i * 6] = left ? (J[i * 6] + diss) : (J[i * 6] - diss);
            if (nm.viscous) {
                viscous_jacobian_dense(vm, Jv);
                for (int i = 0; i < 25; ++i) J[i] += left ? Jv[i] : -Jv[i];
            }
            if (pass == 0) {
                if (own)
                    for (int i = 0; i < 25; ++i) dg[i] += J[i];
                else
                    for (int i = 0; i < 25; ++i) dg[i] -= J[i];
            } else {
                const int k = find_col(b, c, own ? c1 : c0);
                double* dst = values + static_cast<long long>(k) * 25;
                if (own)
                    for (int i = 0; i < 25; ++i) dst[i] = 0.0 + J[i];
                else
                    for (int i = 0; i < 25; ++i) dst[i] = 0.0 - J[i];
            }
        }
    }
    const double a = alpha[c];
    for (int i = 0; i < 5; ++i) dg[i * 6] += a;
    double* dst = values + static_cast<long long>(b.diag[c]) * 25;
    for (int i = 0; i < 25; ++i) dst[i] = dg[i];
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

void launch_gradient_limiter(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, hipStream_t s) {
    if (m.nc > 0) hipLaunchKernelGGL(k_gradient_limiter, grid_for(m.nc), dim3(kBlock), 0, s, m, f, nm);
}

void launch_ghost_gradients(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, const BcTable& bcs,
                            hipStream_t s) {
    const int nbf = m.nf - m.ninner;
    if (nbf > 0) hipLaunchKernelGGL(k_ghost_gradients, grid_for(nbf), dim3(kBlock), 0, s, m, f, nm, bcs);
}

void launch_flux(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, hipStream_t s) {
    if (m.nc > 0) hipLaunchKernelGGL(k_flux, grid_for(m.nc), dim3(kBlock), 0, s, m, f, nm);
}

void launch_dt(const MeshArgs& m, const double* lam, double cfl, double* dt, double* alpha, hipStream_t s) {
    if (m.nc > 0) hipLaunchKernelGGL(k_dt, grid_for(m.nc), dim3(kBlock), 0, s, m.nc, m.vol, lam, cfl, dt, alpha);
}

void launch_jacobian(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, const BsrArgs& b,
                     const double* alpha, double* values, hipStream_t s) {
    if (m.nc > 0) hipLaunchKernelGGL(k_jacobian, grid_for(m.nc), dim3(kBlock), 0, s, m, f, nm, b, alpha, values);
}

}  // namespace cfdg
