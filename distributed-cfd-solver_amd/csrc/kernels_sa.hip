// Spalart-Allmaras module kernels (gfx950), device cell order.
//
//   k_sa_prepare   density, nu_tilde and eddy viscosity of the cells; zone
//                  ghost values of the boundary faces (one thread per cell or
//                  boundary face)
//   k_sa_gradient  nu_tilde gradient with the configured method, one thread
//                  per cell (node least squares with the cell's inverse normal
//                  matrix, as the mean-flow gradient kernel, or Green-Gauss)
//   k_sa_residual  per cell, over its faces in ascending face index: upwind
//                  convection on the recorded mean mass flux, skew-corrected
//                  diffusion, the interior faces' spectral radius added to
//                  lam; then the production / destruction / cb2 sources and
//                  the point-implicit diagonal
//   k_sa_update    frozen (positivity clamp only) or point-implicit update
//
// The equations follow spalart_allmaras.hpp:196-442 with four corrections:
// the diffusive flux leaves the owner as -coef (grad nu . n) A, cb2 is a
// source, the freestream override of a boundary face with inflowing mass is
// limited to inflow / outflow zones, and fw has the (g^6 + cw3^6) denominator
// of the header's formula (spalart_allmaras.hpp:17).  Single writer per output, fixed
// summation order, no atomics.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace cfdg {

namespace {

constexpr int kBlock = 256;
constexpr double kCb1 = 0.1355, kCb2 = 0.622, kSigma = 2.0 / 3.0, kKappa = 0.41;
constexpr double kCw2 = 0.3, kCw3 = 2.0, kCv1 = 7.1;
constexpr double kCw1 = kCb1 / (kKappa * kKappa) + (1.0 + kCb2) / kSigma;
constexpr double kCv1_3 = kCv1 * kCv1 * kCv1;
constexpr double kSbarFloor = 1.0e-10, kMinWallDist = 1.0e-10;

__device__ __forceinline__ CellRef crec(const FieldArgs& f, int c) { return CellRef{f.cs, f.ld, c}; }
__device__ __forceinline__ FaceRef frec(const MeshArgs& m, int f) { return FaceRef{m.fp, m.fld, f}; }

__device__ __forceinline__ double fv1_of(double chi) {
    const double chi3 = chi * chi * chi;
    return chi3 / (chi3 + kCv1_3);
}

__global__ void __launch_bounds__(kBlock) k_sa_prepare(MeshArgs m, FieldArgs f, NumArgs nm, SaArgs sa) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nbf = m.nf - m.ninner;
    if (i >= m.nc + nbf) return;
    const CellRef r = crec(f, i);
    const double rho = r[kRQ] / (nm.gas.R * r[kRQ + 4]);  // density_Tp
    if (i < m.nc) {
        const double nu = sa.rnu[i] / rho;
        sa.rho[i] = rho;
        sa.nu[i] = nu;
        const double nu_lam = sutherland(r[kRQ + 4]) / rho;
        sa.mut[i] = rho * nu * fv1_of(nu / nu_lam);
        return;
    }
    const int b = i - m.nc;  // boundary face ninner + b, ghost slot i
    const int c0 = m.fcells[m.ninner + b].x;
    double rg;
    switch (sa.kind[b]) {
        case kZoneWall: rg = -sa.rnu[c0]; break;
        case kZoneInflow: rg = rho * sa.nu_inf; break;
        default: rg = sa.rnu[c0]; break;
    }
    sa.rnu[i] = rg;
    sa.rho[i] = rho;
    sa.nu[i] = rg / rho;
    sa.mut[i] = 0.0;  // BC ghosts carry no turbulence (spalart_allmaras.hpp:266-268)
}

__global__ void __launch_bounds__(kBlock) k_sa_gradient(MeshArgs m, FieldArgs f, NumArgs nm, SaArgs sa) {
    const int c = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (c >= m.nc) return;
    const CellRef rc = crec(f, c);
    const double s0 = sa.nu[c];
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    if (nm.gradient == 2) {
        const double x0 = rc[kRC], y0 = rc[kRC + 1], z0 = rc[kRC + 2];
        const int jb = m.ccn_off[c], je = m.ccn_off[c + 1];
        if (je > jb) {
            // inverse normal matrix of the cell, set up with the mesh (planes kRInv..)
            const double i0 = rc[kRInv], i1 = rc[kRInv + 1], i2 = rc[kRInv + 2], i4 = rc[kRInv + 3],
                         i5 = rc[kRInv + 4], i8 = rc[kRInv + 5];
            for (int j = jb; j < je; ++j) {
                const int n = m.ccn[j];
                const CellRef rn = crec(f, n);
                const double dx = rn[kRC] - x0, dy = rn[kRC + 1] - y0, dz = rn[kRC + 2] - z0;
                const double w = 1.0 / dmax(dx * dx + dy * dy + dz * dz, 1.0e-30);
                const double rx = w * dx, ry = w * dy, rz = w * dz;
                const double d = sa.nu[n] - s0;
                g0 += (i0 * rx + i1 * ry + i2 * rz) * d;
                g1 += (i1 * rx + i4 * ry + i5 * rz) * d;
                g2 += (i2 * rx + i5 * ry + i8 * rz) * d;
            }
        }
    } else {
        for (int e = m.cf_off[c]; e < m.cf_off[c + 1]; ++e) {
            const int fc = m.cf[e];
            const FaceRef fr = frec(m, fc);
            const int2 oc = m.fcells[fc];
            const double a05 = 0.5 * fr[kFA];
            const double s = sa.nu[oc.x] + sa.nu[oc.y];
            const double fx = a05 * fr[kFN] * s, fy = a05 * fr[kFN + 1] * s, fz = a05 * fr[kFN + 2] * s;
            if (oc.x == c) {
                g0 += fx;
                g1 += fy;
                g2 += fz;
            } else {
                g0 -= fx;
                g1 -= fy;
                g2 -= fz;
            }
        }
        const double iv = 1.0 / rc[kRVol];
        g0 *= iv;
        g1 *= iv;
        g2 *= iv;
    }
    sa.g[c] = g0;
    sa.g[static_cast<long long>(m.nc) + c] = g1;
    sa.g[2LL * m.nc + c] = g2;
}

// Tiled like the mean-flow flux: one workgroup per flux tile.  The tile's
// cells (nu_tilde, density, T, nu_tilde gradient, centroid) are staged in LDS;
// every face touching the tile is evaluated once (its convective + diffusive
// flux, the two diagonal numerators and the spectral-radius term depend on the
// face alone) into LDS; then each cell adds its faces in ascending face index
// -- the order and operations of the per-cell loop of
// spalart_allmaras.hpp:196-380, dividing the diagonal numerators by its own
// density -- and its sources.
constexpr int kSaCellVals = 9;  // nu, rho, T, grad nu (3), centroid (3)
__global__ void __launch_bounds__(kTileCells) k_sa_residual(MeshArgs m, FieldArgs f, NumArgs nm, SaArgs sa) {
    __shared__ double scl[kSaCellVals * kTileCells];
    __shared__ double sfc[4 * kTileFaces];  // flux, owner / neighbour diagonal numerators, lam term
    __shared__ unsigned char sbd[kTileFaces];  // 1: boundary face (no lam term)
    constexpr double inv_sigma = 1.0 / kSigma;
    const int t = xcd_block(blockIdx.x, gridDim.x);
    const int c0 = m.tile_off[t], nt = m.tile_off[t + 1] - c0;
    const int fb = m.tile_face_off[t], nfl = m.tile_face_off[t + 1] - fb;
    const int i = threadIdx.x;
    const long long nc = m.nc;
    if (i < nt) {
        const int c = c0 + i;
        const CellRef rc = crec(f, c);
        scl[0 * kTileCells + i] = sa.nu[c];
        scl[1 * kTileCells + i] = sa.rho[c];
        scl[2 * kTileCells + i] = rc[kRQ + 4];
        scl[3 * kTileCells + i] = sa.g[c];
        scl[4 * kTileCells + i] = sa.g[nc + c];
        scl[5 * kTileCells + i] = sa.g[2 * nc + c];
        scl[6 * kTileCells + i] = rc[kRC];
        scl[7 * kTileCells + i] = rc[kRC + 1];
        scl[8 * kTileCells + i] = rc[kRC + 2];
    }
    __syncthreads();
    // value v of cell x: staged when x lies in the tile
    const auto cellv = [&](int x, bool in, int v) -> double {
        if (in) return scl[v * kTileCells + (x - c0)];
        switch (v) {
            case 0: return sa.nu[x];
            case 1: return sa.rho[x];
            case 2: return crec(f, x)[kRQ + 4];
            case 3: return sa.g[x];
            case 4: return sa.g[nc + x];
            case 5: return sa.g[2 * nc + x];
            case 6: return crec(f, x)[kRC];
            case 7: return crec(f, x)[kRC + 1];
            default: return crec(f, x)[kRC + 2];
        }
    };
    typedef double pair_t __attribute__((ext_vector_type(2)));
    for (int j = i; j < nfl; j += kTileCells) {
        const long long e = fb + j;
        const pair_t w01 = __builtin_nontemporal_load(reinterpret_cast<const pair_t*>(m.tf + plane_at(0, m.tld, e)));
        const pair_t w23 = __builtin_nontemporal_load(reinterpret_cast<const pair_t*>(m.tf + plane_at(2, m.tld, e)));
        const pair_t w45 = __builtin_nontemporal_load(reinterpret_cast<const pair_t*>(m.tf + plane_at(4, m.tld, e)));
        const pair_t w67 = __builtin_nontemporal_load(reinterpret_cast<const pair_t*>(m.tf + plane_at(6, m.tld, e)));
        const double nx = w01.x, ny = w01.y, nz = w23.x, A = w23.y;
        const unsigned long long ids = static_cast<unsigned long long>(__double_as_longlong(w67.y));
        const int o = static_cast<int>(static_cast<unsigned>(ids)), n = static_cast<int>(static_cast<unsigned>(ids >> 32));
        const bool inner = n < m.nc;  // boundary faces: n is the ghost cells + boundary index
        const bool in0 = static_cast<unsigned>(o - c0) < static_cast<unsigned>(nt);
        const bool in1 = static_cast<unsigned>(n - c0) < static_cast<unsigned>(nt);
        const double mf = sa.mdot_tile[e];
        const double nu0 = cellv(o, in0, 0);
        double nu_up, nu1, gb[3];
        const double ga[3] = {cellv(o, in0, 3), cellv(o, in0, 4), cellv(o, in0, 5)};
        const double x0 = cellv(o, in0, 6), y0 = cellv(o, in0, 7), z0 = cellv(o, in0, 8);
        double dx, dy, dz, inv_d;
        if (inner) {
            const double nu1_cell = cellv(n, in1, 0);
            nu_up = (mf >= 0.0) ? nu0 : nu1_cell;
            nu1 = nu1_cell;
            gb[0] = cellv(n, in1, 3);
            gb[1] = cellv(n, in1, 4);
            gb[2] = cellv(n, in1, 5);
            // owner -> neighbour centroid vector (face_distances)
            const double ex = cellv(n, in1, 6) - x0, ey = cellv(n, in1, 7) - y0, ez = cellv(n, in1, 8) - z0;
            dx = ex;
            dy = ey;
            dz = ez;
            inv_d = 1.0 / sqrt(ex * ex + ey * ey + ez * ez);
        } else {
            const int kind = sa.kind[n - m.nc];
            const bool inflow = mf < 0.0 && (kind == kZoneInflow || kind == kZoneOutflow);
            nu_up = inflow ? sa.nu_inf : nu0;
            nu1 = inflow ? sa.nu_inf : sa.rnu[n] / sa.rho[n];
            // the ghost's gradient mirrors the owner's, negated at walls
            const double sg = (kind == kZoneWall) ? -1.0 : 1.0;
            gb[0] = sg * ga[0];
            gb[1] = sg * ga[1];
            gb[2] = sg * ga[2];
            const double ex = w45.x - x0, ey = w45.y - y0, ez = w67.x - z0;
            dx = 2.0 * ex;
            dy = 2.0 * ey;
            dz = 2.0 * ez;
            inv_d = 1.0 / sqrt(4.0 * ex * ex + 4.0 * ey * ey + 4.0 * ez * ez);
        }
        const double f_conv = mf * nu_up;
        // averaged + skew-corrected face gradient (spalart_allmaras.hpp:500-525)
        double gn0 = 0.5 * (ga[0] + gb[0]), gn1 = 0.5 * (ga[1] + gb[1]), gn2 = 0.5 * (ga[2] + gb[2]);
        const double jump = nu1 - nu0;
        const double gd = gn0 * dx + gn1 * dy + gn2 * dz;
        const double corr = (jump - gd) * inv_d * inv_d;
        gn0 += corr * dx;
        gn1 += corr * dy;
        gn2 += corr * dz;
        const double T1 = inner ? cellv(n, in1, 2) : crec(f, n)[kRQ + 4];
        const double rho1 = inner ? cellv(n, in1, 1) : sa.rho[n];
        const double Tf = 0.5 * (cellv(o, in0, 2) + T1);
        const double mu_f = sutherland(Tf);
        const double rho_f = 0.5 * (cellv(o, in0, 1) + rho1);
        const double coef = inner ? inv_sigma * (mu_f + rho_f * (0.5 * (nu0 + nu1)))
                                  : inv_sigma * (mu_f + rho_f * 0.5 * (nu0 + nu1));
        const double f_diff = coef * (gn0 * nx + gn1 * ny + gn2 * nz) * A;
        const double dcoef = coef * A * inv_d;
        sfc[0 * kTileFaces + j] = f_conv - f_diff;
        sfc[1 * kTileFaces + j] = dmax(mf, 0.0) + dcoef;
        sfc[2 * kTileFaces + j] = dmax(-mf, 0.0) + dcoef;
        double lam_add = 0.0;
        if (inner) {
            const double visc_term = 2.0 * inv_sigma * (mu_f + rho_f * (0.5 * (nu0 + nu1))) / rho_f * A * inv_d;
            lam_add = fabs(mf) / rho_f + visc_term;
        }
        sfc[3 * kTileFaces + j] = lam_add;
        sbd[j] = inner ? 0 : 1;
    }
    __syncthreads();
    if (i >= nt) return;
    const int c = c0 + i;
    const CellRef rc = crec(f, c);
    const double rho_c = scl[1 * kTileCells + i];
    double res = 0.0, diag = 0.0;
    double lam = f.lam[c];
    for (int ee = m.cf_off[c]; ee < m.cf_off[c + 1]; ++ee) {
        const int sl = m.cf_slot[ee];
        const int k = (sl >= 0) ? sl : -sl - 1;
        if (!sbd[k]) lam += sfc[3 * kTileFaces + k];  // interior faces only
        if (sl >= 0) {
            res += sfc[0 * kTileFaces + k];
            diag += sfc[1 * kTileFaces + k] / rho_c;
        } else {
            res -= sfc[0 * kTileFaces + k];
            diag += sfc[2 * kTileFaces + k] / rho_c;
        }
    }
    // sources (spalart_allmaras.hpp:382-432)
    const double nu = scl[0 * kTileCells + i];
    const double d = dmax(sa.dist[c], kMinWallDist);
    const double inv_d2 = 1.0 / (d * d);
    const double inv_k2 = 1.0 / (kKappa * kKappa);
    const double ux = rc[kRG + 3], uy = rc[kRG + 4], uz = rc[kRG + 5];
    const double vx = rc[kRG + 6], vy = rc[kRG + 7], vz = rc[kRG + 8];
    const double wx = rc[kRG + 9], wy = rc[kRG + 10], wz = rc[kRG + 11];
    const double s_xy = uy + vx, s_xz = uz + wx, s_yz = vz + wy;
    const double s2 = 2.0 * (ux * ux + vy * vy + wz * wz) + 0.5 * (s_xy * s_xy + s_xz * s_xz + s_yz * s_yz);
    const double S = sqrt(dmax(s2, 0.0));
    const double nu_lam = sutherland(scl[2 * kTileCells + i]) / rho_c;
    const double chi = nu / nu_lam;
    const double fv1 = fv1_of(chi);
    const double fv2 = 1.0 - chi / (1.0 + chi * fv1);
    const double sbar = dmax(S + nu * fv2 * inv_k2 * inv_d2, kSbarFloor);
    double r = nu / (sbar * kKappa * kKappa * d * d);
    r = dmin(dmax(r, 0.0), 10.0);
    const double r2 = r * r;
    const double r6 = r2 * r2 * r2;
    const double g = r + kCw2 * (r6 - r);
    const double g2 = dmax(g * g, kSbarFloor);
    const double g6 = g2 * g2 * g2;
    constexpr double cw3_6 = kCw3 * kCw3 * kCw3 * kCw3 * kCw3 * kCw3;
    const double fw = g * sqrt(cbrt((1.0 + cw3_6) / (g6 + cw3_6)));
    const double P = kCb1 * sbar * rho_c * nu;
    const double D = kCw1 * fw * rho_c * nu * nu * inv_d2;
    const double gx = scl[3 * kTileCells + i], gy = scl[4 * kTileCells + i], gz = scl[5 * kTileCells + i];
    const double cb2 = (kCb2 / kSigma) * rho_c * (gx * gx + gy * gy + gz * gz);
    const double vol = rc[kRVol];
    res += vol * (D - P - cb2);
    const double lin = 2.0 * kCw1 * fw * nu / (d * d) - kCb1 * sbar;
    diag += vol * dmax(lin, 0.0);
    sa.res[c] = res;
    sa.diag[c] = diag;
    f.lam[c] = lam;
}

__global__ void __launch_bounds__(kBlock) k_sa_update(int nc, int point_implicit, const double* alpha, SaArgs sa) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double v = sa.rnu[c];
    if (point_implicit) v = v - sa.res[c] / (alpha[c] + sa.diag[c]);
    sa.rnu[c] = (v < 0.0) ? 0.0 : v;
}

int blocks(long long n) { return static_cast<int>((n + kBlock - 1) / kBlock); }

}  // namespace

void launch_sa_prepare(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, const SaArgs& sa, hipStream_t s) {
    const long long n = static_cast<long long>(m.nc) + (m.nf - m.ninner);
    k_sa_prepare<<<blocks(n), kBlock, 0, s>>>(m, f, nm, sa);
}

void launch_sa_gradient(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, const SaArgs& sa, hipStream_t s) {
    k_sa_gradient<<<blocks(m.nc), kBlock, 0, s>>>(m, f, nm, sa);
}

void launch_sa_residual(const MeshArgs& m, const FieldArgs& f, const NumArgs& nm, const SaArgs& sa, hipStream_t s) {
    k_sa_residual<<<m.ntiles, kTileCells, 0, s>>>(m, f, nm, sa);
}

void launch_sa_update(int nc, int point_implicit, const double* alpha, const SaArgs& sa, hipStream_t s) {
    k_sa_update<<<blocks(nc), kBlock, 0, s>>>(nc, point_implicit, alpha, sa);
}

}  // namespace cfdg
