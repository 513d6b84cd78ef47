// CPU restatement ("oracle") of the reference solver's single-rank implicit
// iteration.  TEST INFRASTRUCTURE ONLY -- never linked into the product.
//
// Every routine restates one reference routine with the same operation order
// (compiled with -ffp-contract=off) so that results can be compared bit for
// bit; the reference file:line is cited at each function.  Paths are relative
// to the reference tree.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using i32 = std::int32_t;
using std::size_t;

// include/cfd/core/types.hpp:45-55
constexpr double kSuthMu = 1.716e-5;
constexpr double kSuthTref = 273.15;
constexpr double kSuthS = 110.4;
constexpr double kLamFloor = 1.0e-14;
constexpr double kNormFloor = 1.0e-300;

enum Bc { SupIn = 0, SupOut, SubIn, SubOut, Slip, NoSlip, NoSlipHF, Sym, Far };

struct Eos {  // eos/ideal_gas.hpp:13-85
    double g, R;
    double gm1() const { return g - 1.0; }
    double cv() const { return R / (g - 1.0); }
    double cp() const { return g * cv(); }
    double density_Tp(double T, double p) const { return p / (R * T); }
    double internal_energy_Tp(double T) const { return cv() * T; }
    double temperature_rhoe(double e) const { return e / cv(); }
    double temperature_rhop(double rho, double p) const { return p / (rho * R); }
    double pressure_rhoe(double rho, double e) const { return gm1() * rho * e; }
    double sound_speed_rhop(double rho, double p) const { return std::sqrt(g * p / rho); }
    double sound_speed_Hv2(double H, double v2) const {
        const double a2 = gm1() * (H - 0.5 * v2);
        return std::sqrt(a2 > 0.0 ? a2 : 0.0);
    }
};

// state_conversions.hpp:12-17, 21-25
double specific_internal_energy(const double U[5]) {
    const double rho = U[0];
    const double inv_rho = 1.0 / rho;
    const double q2 = (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) * (inv_rho * inv_rho);
    return (U[4] * inv_rho) - 0.5 * q2;
}
double pressure(const Eos& e, const double U[5]) { return e.pressure_rhoe(U[0], specific_internal_energy(U)); }

// state_conversions.hpp:74-89
void cons_to_prim(const Eos& eos, const double U[5], double& p, double& u, double& v, double& w, double& T) {
    const double rho = U[0];
    const double inv_rho = 1.0 / rho;
    u = U[1] * inv_rho;
    v = U[2] * inv_rho;
    w = U[3] * inv_rho;
    const double q2 = u * u + v * v + w * w;
    const double e = (U[4] * inv_rho) - 0.5 * q2;
    p = eos.pressure_rhoe(rho, e);
    T = eos.temperature_rhoe(e);
}

// state_conversions.hpp:115-129
void prim_to_cons(const Eos& eos, double p, double u, double v, double w, double T, double U[5]) {
    const double rho = eos.density_Tp(T, p);
    const double e = eos.internal_energy_Tp(T);
    const double ke = 0.5 * rho * (u * u + v * v + w * w);
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = rho * w;
    U[4] = rho * e + ke;
}

// hllc.hpp:231-274
void star_flux(const double UK[5], const double FK[5], double rhoK, double pK, double ux, double uy, double uz,
               double unk, double ak, double SK, double SM, double nx, double ny, double nz, double area,
               double F[5]) {
    const double den = SK - SM;
    if (std::fabs(den) < 1.0e-6 * (ak + std::fabs(SK))) {
        for (int v = 0; v < 5; ++v) F[v] = FK[v] * area;
        return;
    }
    const double inv_den = 1.0 / den;
    const double d_wave = SK - unk;
    const double q = d_wave * inv_den;
    const double d = SM - unk;
    const double p_term = pK / (rhoK * d_wave);
    const double e_star = UK[4] + rhoK * d * (SM + p_term);
    const double Ust[5] = {q * rhoK, q * rhoK * (ux + d * nx), q * rhoK * (uy + d * ny), q * rhoK * (uz + d * nz),
                           q * e_star};
    for (int v = 0; v < 5; ++v) F[v] = (FK[v] + SK * (Ust[v] - UK[v])) * area;
}

// hllc.hpp:22-122
void hllc(const Eos& eos, const double UL[5], const double UR[5], double nx, double ny, double nz, double area,
          double F[5], double& smax) {
    const double rhoL = UL[0], rhoR = UR[0];
    const double inv_rhoL = 1.0 / rhoL, inv_rhoR = 1.0 / rhoR;
    const double pL = pressure(eos, UL), pR = pressure(eos, UR);
    const double uLx = UL[1] * inv_rhoL, uLy = UL[2] * inv_rhoL, uLz = UL[3] * inv_rhoL;
    const double uRx = UR[1] * inv_rhoR, uRy = UR[2] * inv_rhoR, uRz = UR[3] * inv_rhoR;
    const double unL = uLx * nx + uLy * ny + uLz * nz;
    const double unR = uRx * nx + uRy * ny + uRz * nz;
    const double aL = eos.sound_speed_rhop(rhoL, pL);
    const double aR = eos.sound_speed_rhop(rhoR, pR);
    const double rl = std::sqrt(rhoL), rr = std::sqrt(rhoR);
    const double w = 1.0 / (rl + rr);
    const double roe_ux = (rl * uLx + rr * uRx) * w;
    const double roe_uy = (rl * uLy + rr * uRy) * w;
    const double roe_uz = (rl * uLz + rr * uRz) * w;
    const double HL = (UL[4] + pL) * inv_rhoL;
    const double HR = (UR[4] + pR) * inv_rhoR;
    const double roe_h = (rl * HL + rr * HR) * w;
    const double q2 = roe_ux * roe_ux + roe_uy * roe_uy + roe_uz * roe_uz;
    const double roe_a = eos.sound_speed_Hv2(roe_h, q2);
    const double roe_un = roe_ux * nx + roe_uy * ny + roe_uz * nz;
    const double SL = std::min(unL - aL, roe_un - roe_a);
    const double SR = std::max(unR + aR, roe_un + roe_a);
    smax = std::max(std::fabs(SL), std::fabs(SR));
    const double dL = SL - unL, dR = SR - unR;
    const double denom = rhoL * dL - rhoR * dR;
    const double SM = (std::fabs(denom) > 1e-14) ? (pR - pL + rhoL * unL * dL - rhoR * unR * dR) / denom : 0.5 * (unL + unR);
    if (SM >= 0.0) {
        const double FL[5] = {rhoL * unL, rhoL * uLx * unL + pL * nx, rhoL * uLy * unL + pL * ny,
                              rhoL * uLz * unL + pL * nz, (UL[4] + pL) * unL};
        if (SL >= 0.0) {
            for (int v = 0; v < 5; ++v) F[v] = FL[v] * area;
        } else {
            star_flux(UL, FL, rhoL, pL, uLx, uLy, uLz, unL, aL, SL, SM, nx, ny, nz, area, F);
        }
    } else {
        const double FR[5] = {rhoR * unR, rhoR * uRx * unR + pR * nx, rhoR * uRy * unR + pR * ny,
                              rhoR * uRz * unR + pR * nz, (UR[4] + pR) * unR};
        if (SR <= 0.0) {
            for (int v = 0; v < 5; ++v) F[v] = FR[v] * area;
        } else {
            star_flux(UR, FR, rhoR, pR, uRx, uRy, uRz, unR, aR, SR, SM, nx, ny, nz, area, F);
        }
    }
}

// hllc.hpp:173-229
void euler_jacobian(const Eos& eos, const double U[5], double nx, double ny, double nz, double J[25]) {
    const double rho = U[0];
    const double inv_rho = 1.0 / rho;
    const double u = U[1] * inv_rho, v = U[2] * inv_rho, w = U[3] * inv_rho, E = U[4] * inv_rho;
    const double p = pressure(eos, U);
    const double gamma = eos.g;
    const double phi = gamma - 1.0;
    const double un = u * nx + v * ny + w * nz;
    const double q2 = u * u + v * v + w * w;
    const double H = E + p * inv_rho;
    const double hpq = 0.5 * phi * q2;
    J[0] = 0.0; J[1] = nx; J[2] = ny; J[3] = nz; J[4] = 0.0;
    J[5] = hpq * nx - u * un; J[6] = un + nx * (1.0 - phi) * u; J[7] = u * ny - phi * nx * v;
    J[8] = u * nz - phi * nx * w; J[9] = phi * nx;
    J[10] = hpq * ny - v * un; J[11] = v * nx - phi * ny * u; J[12] = un + ny * (1.0 - phi) * v;
    J[13] = v * nz - phi * ny * w; J[14] = phi * ny;
    J[15] = hpq * nz - w * un; J[16] = w * nx - phi * nz * u; J[17] = w * ny - phi * nz * v;
    J[18] = un + nz * (1.0 - phi) * w; J[19] = phi * nz;
    J[20] = un * (hpq - H); J[21] = H * nx - phi * u * un; J[22] = H * ny - phi * v * un;
    J[23] = H * nz - phi * w * un; J[24] = gamma * un;
}

// hllc.hpp:136-168
void hllc_jacobian(const Eos& eos, const double UL[5], const double UR[5], double nx, double ny, double nz,
                   double area, double dFL[25], double dFR[25]) {
    double F0[5];
    double smax = 0.0;
    hllc(eos, UL, UR, nx, ny, nz, area, F0, smax);
    euler_jacobian(eos, UL, nx, ny, nz, dFL);
    euler_jacobian(eos, UR, nx, ny, nz, dFR);
    const double half_area = 0.5 * area;
    const double diss = 0.5 * smax * area;
    for (int i = 0; i < 25; ++i) {
        dFL[i] *= half_area;
        dFR[i] *= half_area;
    }
    for (int i = 0; i < 5; ++i) {
        dFL[i * 5 + i] += diss;
        dFR[i * 5 + i] -= diss;
    }
}

// viscous.hpp:68-81
double viscosity(double T) {
    constexpr double kSuthConst = kSuthMu * (kSuthTref + kSuthS);
    const double ratio = T / kSuthTref;
    return kSuthConst * ratio * std::sqrt(ratio) / (T + kSuthS);
}
double conductivity(const Eos& eos, double T, double prandtl) { return (viscosity(T) * eos.cp()) / prandtl; }

struct Views {  // viscous.hpp:21-64: primitive values and gradients by cell index
    const double* q[5];
    const double* gx[5];
    const double* gy[5];
    const double* gz[5];
};

// viscous.hpp:90-174 (SoA overload)
void viscous_flux(const Eos& eos, double inv_d, double dvx, double dvy, double dvz, size_t c0, size_t c1,
                  const Views& vv, double mutL, double mutR, double nx, double ny, double nz, double area,
                  double Fv[5], double& lam_visc, double pr_lam, double pr_turb) {
    const double p_f = 0.5 * (vv.q[0][c0] + vv.q[0][c1]);
    const double u_f = 0.5 * (vv.q[1][c0] + vv.q[1][c1]);
    const double v_f = 0.5 * (vv.q[2][c0] + vv.q[2][c1]);
    const double w_f = 0.5 * (vv.q[3][c0] + vv.q[3][c1]);
    const double T_f = 0.5 * (vv.q[4][c0] + vv.q[4][c1]);
    const double mu_lam = viscosity(T_f);
    const double mu_t_f = 0.5 * (mutL + mutR);
    const double mu_eff = mu_lam + mu_t_f;
    const double k_lam = conductivity(eos, T_f, pr_lam);
    const double cp = eos.cp();
    const double k_eff = k_lam + (mu_t_f * cp) / pr_turb;
    const double xi_x = dvx * inv_d, xi_y = dvy * inv_d, xi_z = dvz * inv_d;
    const double ncx = nx - xi_x, ncy = ny - xi_y, ncz = nz - xi_z;
    const auto avg = [&](const double* const* g, int var) { return 0.5 * (g[var][c0] + g[var][c1]); };
    const double du_dx = avg(vv.gx, 1), du_dy = avg(vv.gy, 1), du_dz = avg(vv.gz, 1);
    const double dv_dx = avg(vv.gx, 2), dv_dy = avg(vv.gy, 2), dv_dz = avg(vv.gz, 2);
    const double dw_dx = avg(vv.gx, 3), dw_dy = avg(vv.gy, 3), dw_dz = avg(vv.gz, 3);
    const double dT_dx = avg(vv.gx, 4), dT_dy = avg(vv.gy, 4), dT_dz = avg(vv.gz, 4);
    const double div_v = du_dx + dv_dy + dw_dz;
    const double ttd = (2.0 / 3.0) * div_v;
    const double du_dn = (vv.q[1][c1] - vv.q[1][c0]) * inv_d + (du_dx * ncx + du_dy * ncy + du_dz * ncz);
    const double dv_dn = (vv.q[2][c1] - vv.q[2][c0]) * inv_d + (dv_dx * ncx + dv_dy * ncy + dv_dz * ncz);
    const double dw_dn = (vv.q[3][c1] - vv.q[3][c0]) * inv_d + (dw_dx * ncx + dw_dy * ncy + dw_dz * ncz);
    const double dT_dn = (vv.q[4][c1] - vv.q[4][c0]) * inv_d + (dT_dx * ncx + dT_dy * ncy + dT_dz * ncz);
    const double tnx = mu_eff * (du_dn + (du_dx * nx + dv_dx * ny + dw_dx * nz) - ttd * nx);
    const double tny = mu_eff * (dv_dn + (du_dy * nx + dv_dy * ny + dw_dy * nz) - ttd * ny);
    const double tnz = mu_eff * (dw_dn + (du_dz * nx + dv_dz * ny + dw_dz * nz) - ttd * nz);
    const double qn = -k_eff * dT_dn;
    const double rho_f = eos.density_Tp(T_f, p_f);
    const double nu_eff = mu_eff / rho_f;
    lam_visc = (4.0 / 3.0) * nu_eff * inv_d * area;
    Fv[0] = 0.0;
    Fv[1] = -tnx * area;
    Fv[2] = -tny * area;
    Fv[3] = -tnz * area;
    Fv[4] = (-(u_f * tnx + v_f * tny + w_f * tnz) + qn) * area;
}

// viscous.hpp:280-359
void viscous_jacobian(const Eos& eos, double inv_d, const double qL[5], const double qR[5], double mutL,
                      double mutR, double area, double dFL[25], double dFR[25], double pr_lam, double pr_turb) {
    const double p_f = 0.5 * (qL[0] + qR[0]);
    const double u_f = 0.5 * (qL[1] + qR[1]);
    const double v_f = 0.5 * (qL[2] + qR[2]);
    const double w_f = 0.5 * (qL[3] + qR[3]);
    const double T_f = 0.5 * (qL[4] + qR[4]);
    const double rho_f = eos.density_Tp(T_f, p_f);
    const double inv_rho = 1.0 / rho_f;
    const double cv = eos.cv();
    const double inv_rho_cv = 1.0 / (rho_f * cv);
    const double mu_lam = viscosity(T_f);
    const double mu_t_f = 0.5 * (mutL + mutR);
    const double mu_eff = mu_lam + mu_t_f;
    const double k_lam = conductivity(eos, T_f, pr_lam);
    const double cp = eos.cp();
    const double k_eff = k_lam + (mu_t_f * cp) / pr_turb;
    const double C_mu = mu_eff * inv_d * area;
    const double C_lam = k_eff * inv_d * area;
    const double V2 = u_f * u_f + v_f * v_f + w_f * w_f;
    for (int i = 0; i < 25; ++i) dFL[i] = 0.0;
    dFL[5] = -C_mu * u_f * inv_rho;
    dFL[6] = C_mu * inv_rho;
    dFL[10] = -C_mu * v_f * inv_rho;
    dFL[12] = C_mu * inv_rho;
    dFL[15] = -C_mu * w_f * inv_rho;
    dFL[18] = C_mu * inv_rho;
    dFL[20] = C_lam * (0.5 * V2 - cv * T_f) * inv_rho_cv - C_mu * V2 * inv_rho;
    dFL[21] = -C_lam * u_f * inv_rho_cv + C_mu * u_f * inv_rho;
    dFL[22] = -C_lam * v_f * inv_rho_cv + C_mu * v_f * inv_rho;
    dFL[23] = -C_lam * w_f * inv_rho_cv + C_mu * w_f * inv_rho;
    dFL[24] = C_lam * inv_rho_cv;
    for (int i = 0; i < 25; ++i) dFR[i] = -dFL[i];
}

// slope_limiters.hpp:21-92
double limiter_phi(int kind, double d_nb, double d_face, double eps2) {
    switch (kind) {
        case 0: {  // Venkatakrishnan
            if (d_face == 0.0) return 1.0;
            const double prod = d_nb * d_face;
            if (prod <= 0.0) return 0.0;
            const double a2 = d_nb * d_nb + eps2;
            const double num = a2 + 2.0 * prod;
            const double den = a2 + 2.0 * d_face * d_face + prod;
            return num / den;
        }
        case 1: {  // Barth-Jespersen
            if (d_face == 0.0) return 1.0;
            const double r = d_nb / d_face;
            return std::clamp(r, 0.0, 1.0);
        }
        case 2: {  // Van Albada
            if (d_face == 0.0) return 1.0;
            const double prod = d_nb * d_face;
            if (prod <= 0.0) return 0.0;
            const double num = d_nb * d_nb + prod;
            const double den = d_nb * d_nb + d_face * d_face + eps2;
            return (den > 0.0) ? (num / den) : 1.0;
        }
        default:
            return 1.0;
    }
}

// gradient_operator.cpp:20-47
bool invert3x3(const double A[9], double inv[9]) {
    constexpr double eps = 1.0e-14;
    const double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                       A[2] * (A[3] * A[7] - A[4] * A[6]);
    if (std::abs(det) <= eps) {
        std::fill_n(inv, 9, 0.0);
        return false;
    }
    const double id = 1.0 / det;
    inv[0] = (A[4] * A[8] - A[5] * A[7]) * id;
    inv[1] = (A[2] * A[7] - A[1] * A[8]) * id;
    inv[2] = (A[1] * A[5] - A[2] * A[4]) * id;
    inv[3] = (A[5] * A[6] - A[3] * A[8]) * id;
    inv[4] = (A[0] * A[8] - A[2] * A[6]) * id;
    inv[5] = (A[2] * A[3] - A[0] * A[5]) * id;
    inv[6] = (A[3] * A[7] - A[4] * A[6]) * id;
    inv[7] = (A[1] * A[6] - A[0] * A[7]) * id;
    inv[8] = (A[0] * A[4] - A[1] * A[3]) * id;
    return true;
}

// bsr_helpers.hpp:278-321
bool lu_factor5(double* a, i32* piv) {
    for (int k = 0; k < 5; ++k) {
        int m = k;
        double mx = std::abs(a[k * 5 + k]);
        for (int i = k + 1; i < 5; ++i) {
            const double v = std::abs(a[i * 5 + k]);
            if (v > mx) {
                mx = v;
                m = i;
            }
        }
        piv[k] = m;
        if (mx == 0.0) return false;
        if (m != k) {
            for (int j = 0; j < 5; ++j) std::swap(a[k * 5 + j], a[m * 5 + j]);
        }
        const double inv_d = 1.0 / a[k * 5 + k];
        for (int i = k + 1; i < 5; ++i) {
            const double l = a[i * 5 + k] * inv_d;
            a[i * 5 + k] = l;
            for (int j = k + 1; j < 5; ++j) a[i * 5 + j] -= l * a[k * 5 + j];
        }
    }
    return true;
}

// bsr_helpers.hpp:325-353
void lu_solve5(const double* lu, const i32* piv, double* b) {
    for (int k = 0; k < 5; ++k) {
        if (piv[k] != k) std::swap(b[k], b[piv[k]]);
    }
    for (int i = 1; i < 5; ++i) {
        double s = b[i];
        for (int j = 0; j < i; ++j) s -= lu[i * 5 + j] * b[j];
        b[i] = s;
    }
    for (int i = 4; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < 5; ++j) s -= lu[i * 5 + j] * b[j];
        b[i] = s / lu[i * 5 + i];
    }
}

struct BcParams {  // resolved per-patch parameters (bc_manager.hpp:235-352)
    int type = Sym;
    double p = 0, u = 0, v = 0, w = 0, T = 0;  // inlet / farfield / wall
    double tmp_grad = 0;
    i32 fbeg = 0, fend = 0;
};

}  // namespace

struct oracle_ctx {
    std::string err;
    // mesh
    i32 nc = 0, nf = 0, ninner = 0, nbf = 0, ntot = 0, npatch = 0, n_nodes = 0;
    std::vector<int> patch_type;  // boundary condition type of each patch
    std::vector<i32> owner, neigh, patch, cell_nodes_off, cell_nodes;
    std::vector<double> cx, cy, cz, vol, fx, fy, fz, nx, ny, nz, area;
    // numerics
    oracle_numerics num{};
    Eos eos{1.4, 287.052874};
    std::vector<BcParams> bcs;
    // aux connectivity (aux_connectivity.cpp)
    std::vector<i32> cf_off, cf, ccf_off, ccf, ccn_off, ccn;
    std::vector<double> lsq_w;  // [3][E]
    std::vector<double> dvx, dvy, dvz, dinv;
    // fields (SoA, n_total)
    std::vector<double> U, q, gx, gy, gz, phi, res, lam, dt, alpha;
    std::vector<double> fscale;  // sum over the cell's faces of |F|, per variable
    std::vector<double> bflux;   // F (inviscid + viscous) of every boundary face, 5 per face
    bool have_residual = false;
    // BSR
    std::vector<i32> row_ptr, cols, diag;
    std::vector<double> values;
    std::vector<double> du;
    // Spalart-Allmaras module (spalart_allmaras.hpp:61-544); off when model == 0
    struct {
        oracle_turbulence cfg{};
        double nu_inf = 0.0, rho_nu_inf = 0.0;
        std::vector<int> kind;  // zone treatment per boundary face (NuZoneKind)
        std::vector<double> dist, rnu, nu, rho, mut, g, res, mdot, diag;
    } sa;
    bool sa_on() const { return sa.cfg.model == 1; }
    // prescribed eddy viscosity (cells + ghosts), the mut argument of
    // ResidualKernel::apply / JacobianKernel::apply; empty when not set
    std::vector<double> mut_ext;
    // eddy viscosity of cell or ghost i as the mean-flow flux sees it
    double mut_at(size_t i) const {
        if (sa_on()) return sa.mut[i];
        return mut_ext.empty() ? 0.0 : mut_ext[i];
    }
    // dual time stepping (solver.hpp:355-380): BDF order 0 = steady, else the
    // physical time step and the stored time levels U^n, U^{n-1} (5 planes of nc)
    int bdf_order = 0, bdf_cur = 1;
    int64_t phys_step = 0;  // physical steps taken since dual time was set
    double phys_dt = 0.0;
    std::vector<double> Un, Unm1;
};

namespace {

double* plane(std::vector<double>& a, i32 n, int v) { return a.data() + static_cast<size_t>(v) * static_cast<size_t>(n); }

// aux_connectivity.cpp:11-86 -- faces of each cell in ascending face order
void build_cell_faces(oracle_ctx& c) {
    c.cf_off.assign(static_cast<size_t>(c.nc) + 1, 0);
    for (i32 f = 0; f < c.nf; ++f) {
        ++c.cf_off[static_cast<size_t>(c.owner[static_cast<size_t>(f)]) + 1];
        if (c.neigh[static_cast<size_t>(f)] >= 0) ++c.cf_off[static_cast<size_t>(c.neigh[static_cast<size_t>(f)]) + 1];
    }
    for (i32 i = 0; i < c.nc; ++i) c.cf_off[static_cast<size_t>(i) + 1] += c.cf_off[static_cast<size_t>(i)];
    c.cf.resize(static_cast<size_t>(c.cf_off.back()));
    std::vector<i32> cur(c.cf_off.begin(), c.cf_off.end() - 1);
    for (i32 f = 0; f < c.nf; ++f) {
        c.cf[static_cast<size_t>(cur[static_cast<size_t>(c.owner[static_cast<size_t>(f)])]++)] = f;
        if (c.neigh[static_cast<size_t>(f)] >= 0) c.cf[static_cast<size_t>(cur[static_cast<size_t>(c.neigh[static_cast<size_t>(f)])]++)] = f;
    }
}

// aux_connectivity.cpp:88-187 -- face neighbours of each cell, sorted
void build_cell_cells_face(oracle_ctx& c) {
    c.ccf_off.assign(static_cast<size_t>(c.nc) + 1, 0);
    for (i32 f = 0; f < c.ninner; ++f) {
        ++c.ccf_off[static_cast<size_t>(c.owner[static_cast<size_t>(f)]) + 1];
        ++c.ccf_off[static_cast<size_t>(c.neigh[static_cast<size_t>(f)]) + 1];
    }
    for (i32 i = 0; i < c.nc; ++i) c.ccf_off[static_cast<size_t>(i) + 1] += c.ccf_off[static_cast<size_t>(i)];
    c.ccf.resize(static_cast<size_t>(c.ccf_off.back()));
    std::vector<i32> cur(c.ccf_off.begin(), c.ccf_off.end() - 1);
    for (i32 f = 0; f < c.ninner; ++f) {
        const i32 o = c.owner[static_cast<size_t>(f)], n = c.neigh[static_cast<size_t>(f)];
        c.ccf[static_cast<size_t>(cur[static_cast<size_t>(o)]++)] = n;
        c.ccf[static_cast<size_t>(cur[static_cast<size_t>(n)]++)] = o;
    }
    for (i32 i = 0; i < c.nc; ++i)
        std::sort(c.ccf.begin() + c.ccf_off[static_cast<size_t>(i)], c.ccf.begin() + c.ccf_off[static_cast<size_t>(i) + 1]);
}

// aux_connectivity.cpp:189-262, 348-445 -- vertex-sharing neighbours, sorted
void build_cell_cells_node(oracle_ctx& c, i32 n_nodes) {
    std::vector<i32> nc_off(static_cast<size_t>(n_nodes) + 1, 0), ncl;
    for (i32 cell = 0; cell < c.nc; ++cell)
        for (i32 k = c.cell_nodes_off[static_cast<size_t>(cell)]; k < c.cell_nodes_off[static_cast<size_t>(cell) + 1]; ++k)
            ++nc_off[static_cast<size_t>(c.cell_nodes[static_cast<size_t>(k)]) + 1];
    for (i32 n = 0; n < n_nodes; ++n) nc_off[static_cast<size_t>(n) + 1] += nc_off[static_cast<size_t>(n)];
    ncl.resize(static_cast<size_t>(nc_off.back()));
    std::vector<i32> cur(nc_off.begin(), nc_off.end() - 1);
    for (i32 cell = 0; cell < c.nc; ++cell)
        for (i32 k = c.cell_nodes_off[static_cast<size_t>(cell)]; k < c.cell_nodes_off[static_cast<size_t>(cell) + 1]; ++k)
            ncl[static_cast<size_t>(cur[static_cast<size_t>(c.cell_nodes[static_cast<size_t>(k)])]++)] = cell;

    std::vector<i32> stamp(static_cast<size_t>(c.nc), -1);
    c.ccn_off.assign(static_cast<size_t>(c.nc) + 1, 0);
    c.ccn.clear();
    for (i32 cell = 0; cell < c.nc; ++cell) {
        const size_t start = c.ccn.size();
        for (i32 k = c.cell_nodes_off[static_cast<size_t>(cell)]; k < c.cell_nodes_off[static_cast<size_t>(cell) + 1]; ++k) {
            const i32 n = c.cell_nodes[static_cast<size_t>(k)];
            for (i32 j = nc_off[static_cast<size_t>(n)]; j < nc_off[static_cast<size_t>(n) + 1]; ++j) {
                const i32 cand = ncl[static_cast<size_t>(j)];
                if (cand == cell || stamp[static_cast<size_t>(cand)] == cell) continue;
                stamp[static_cast<size_t>(cand)] = cell;
                c.ccn.push_back(cand);
            }
        }
        std::sort(c.ccn.begin() + static_cast<std::ptrdiff_t>(start), c.ccn.end());
        c.ccn_off[static_cast<size_t>(cell) + 1] = static_cast<i32>(c.ccn.size());
    }
}

// gradient_operator.cpp:622-707 -- LSQ coefficients C_j = A^-1 w_j dr_j
void setup_lsq(oracle_ctx& c) {
    const size_t E = c.ccn.size();
    c.lsq_w.assign(3 * E, 0.0);
    double* wx = c.lsq_w.data();
    double* wy = wx + E;
    double* wz = wy + E;
    std::vector<double> dr, wb;
    for (i32 cell = 0; cell < c.nc; ++cell) {
        const size_t p1 = static_cast<size_t>(c.ccn_off[static_cast<size_t>(cell)]);
        const size_t p2 = static_cast<size_t>(c.ccn_off[static_cast<size_t>(cell) + 1]);
        if (p2 == p1) continue;
        const double x0 = c.cx[static_cast<size_t>(cell)], y0 = c.cy[static_cast<size_t>(cell)], z0 = c.cz[static_cast<size_t>(cell)];
        dr.assign(3 * (p2 - p1), 0.0);
        wb.assign(p2 - p1, 0.0);
        double A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (size_t i = 0; i < p2 - p1; ++i) {
            const size_t nb = static_cast<size_t>(c.ccn[p1 + i]);
            const double dx = c.cx[nb] - x0, dy = c.cy[nb] - y0, dz = c.cz[nb] - z0;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double w = 1.0 / std::max(r2, 1.0e-30);
            dr[3 * i] = dx;
            dr[3 * i + 1] = dy;
            dr[3 * i + 2] = dz;
            wb[i] = w;
            A[0] += w * dx * dx;
            A[1] += w * dx * dy;
            A[2] += w * dx * dz;
            A[4] += w * dy * dy;
            A[5] += w * dy * dz;
            A[8] += w * dz * dz;
        }
        A[3] = A[1];
        A[6] = A[2];
        A[7] = A[5];
        A[0] += 1.0e-14;
        A[4] += 1.0e-14;
        A[8] += 1.0e-14;
        double inv[9];
        invert3x3(A, inv);
        for (size_t i = 0; i < p2 - p1; ++i) {
            const double w = wb[i];
            const double rx = w * dr[3 * i], ry = w * dr[3 * i + 1], rz = w * dr[3 * i + 2];
            wx[p1 + i] = inv[0] * rx + inv[1] * ry + inv[2] * rz;
            wy[p1 + i] = inv[3] * rx + inv[4] * ry + inv[5] * rz;
            wz[p1 + i] = inv[6] * rx + inv[7] * ry + inv[8] * rz;
        }
    }
}

// aux_geometry.cpp:11-51 (distance vectors) and 120-156 (direct inverse distance)
void setup_face_distances(oracle_ctx& c) {
    const size_t nf = static_cast<size_t>(c.nf);
    c.dvx.resize(nf);
    c.dvy.resize(nf);
    c.dvz.resize(nf);
    c.dinv.resize(nf);
    for (size_t f = 0; f < nf; ++f) {
        const size_t u = static_cast<size_t>(c.owner[f]);
        if (static_cast<i32>(f) < c.ninner) {
            const size_t v = static_cast<size_t>(c.neigh[f]);
            c.dvx[f] = c.cx[v] - c.cx[u];
            c.dvy[f] = c.cy[v] - c.cy[u];
            c.dvz[f] = c.cz[v] - c.cz[u];
            const double d = std::sqrt((c.cx[v] - c.cx[u]) * (c.cx[v] - c.cx[u]) + (c.cy[v] - c.cy[u]) * (c.cy[v] - c.cy[u]) +
                                       (c.cz[v] - c.cz[u]) * (c.cz[v] - c.cz[u]));
            c.dinv[f] = 1.0 / d;
        } else {
            c.dvx[f] = 2.0 * (c.fx[f] - c.cx[u]);
            c.dvy[f] = 2.0 * (c.fy[f] - c.cy[u]);
            c.dvz[f] = 2.0 * (c.fz[f] - c.cz[u]);
            const double d = std::sqrt(4.0 * (c.fx[f] - c.cx[u]) * (c.fx[f] - c.cx[u]) +
                                       4.0 * (c.fy[f] - c.cy[u]) * (c.fy[f] - c.cy[u]) +
                                       4.0 * (c.fz[f] - c.cz[u]) * (c.fz[f] - c.cz[u]));
            c.dinv[f] = 1.0 / d;
        }
    }
}

// bsr_helpers.hpp:16-59 -- row i = sort({i} U face neighbours), diag position
void build_bsr_pattern(oracle_ctx& c) {
    const size_t n = static_cast<size_t>(c.nc);
    c.row_ptr.assign(n + 1, 0);
    c.diag.resize(n);
    for (size_t i = 0; i < n; ++i) c.row_ptr[i + 1] = c.row_ptr[i] + 1 + (c.ccf_off[i + 1] - c.ccf_off[i]);
    c.cols.resize(static_cast<size_t>(c.row_ptr[n]));
    for (size_t i = 0; i < n; ++i) {
        size_t w = static_cast<size_t>(c.row_ptr[i]);
        const size_t start = w;
        c.cols[w++] = static_cast<i32>(i);
        for (i32 e = c.ccf_off[i]; e < c.ccf_off[i + 1]; ++e) c.cols[w++] = c.ccf[static_cast<size_t>(e)];
        std::sort(c.cols.begin() + static_cast<std::ptrdiff_t>(start), c.cols.begin() + static_cast<std::ptrdiff_t>(w));
        c.diag[i] = static_cast<i32>(std::lower_bound(c.cols.begin() + static_cast<std::ptrdiff_t>(start),
                                                      c.cols.begin() + static_cast<std::ptrdiff_t>(w), static_cast<i32>(i)) -
                                     c.cols.begin());
    }
    c.values.assign(static_cast<size_t>(c.row_ptr[n]) * 25, 0.0);
}

// jacobian_kernel.hpp:22-36
size_t bsr_pos(const oracle_ctx& c, size_t row, i32 col) {
    const auto b = c.cols.begin() + c.row_ptr[row];
    const auto e = c.cols.begin() + c.row_ptr[row + 1];
    const auto it = std::lower_bound(b, e, col);
    if (it == e || *it != col) throw std::runtime_error("bsr: element missing in dual graph");
    return static_cast<size_t>(it - c.cols.begin());
}

// bc_manager.hpp:37-111 -- inflow states
void resolve_inflow(const Eos& eos, const oracle_bc& d, double p_ref, double& p, double& u, double& v, double& w,
                    double& T) {
    p = p_ref;
    T = d.t;
    if (d.inflow_mode == 0) {
        u = d.velocity[0];
        v = d.velocity[1];
        w = d.velocity[2];
    } else if (d.inflow_mode == 1) {
        const double kDegToRad = M_PI / 180.0;
        const double ar = d.alpha_deg * kDegToRad, br = d.beta_deg * kDegToRad;
        const double rho = eos.density_Tp(d.t, p_ref);
        const double a = eos.sound_speed_rhop(rho, p_ref);
        const double vm = d.mach * a;
        u = vm * std::cos(ar) * std::cos(br);
        v = -vm * std::sin(br);
        w = vm * std::sin(ar) * std::cos(br);
    } else {
        const double rho = eos.density_Tp(d.t, p_ref);
        const double a = eos.sound_speed_rhop(rho, p_ref);
        const double vm = d.mach * a;
        const double norm = std::sqrt(d.direction[0] * d.direction[0] + d.direction[1] * d.direction[1] +
                                      d.direction[2] * d.direction[2]);
        const double inv = (norm > 1.0e-14) ? (1.0 / norm) : 0.0;
        u = vm * d.direction[0] * inv;
        v = vm * d.direction[1] * inv;
        w = vm * d.direction[2] * inv;
    }
}

// farfield.hpp:39-113
void farfield_state(const BcParams& p, double gamma, double R, double p_in, double T_in, double vx, double vy,
                    double vz, double nx, double ny, double nz, double& pb, double& Tb, double& vxb, double& vyb,
                    double& vzb, bool& sup_out) {
    const double gm1 = gamma - 1.0;
    const double inv_gm1 = 1.0 / gm1;
    const double un_in = vx * nx + vy * ny + vz * nz;
    const double un_inf = p.u * nx + p.v * ny + p.w * nz;
    const double a_in = std::sqrt(gamma * R * std::max(T_in, 1.0e-6));
    const double a_inf = std::sqrt(gamma * R * std::max(p.T, 1.0e-6));
    const double mn_in = un_in / a_in;
    if (mn_in >= 1.0) {
        sup_out = true;
        pb = p_in;
        Tb = T_in;
        vxb = vx;
        vyb = vy;
        vzb = vz;
        return;
    }
    sup_out = false;
    if (mn_in <= -1.0) {
        pb = p.p;
        Tb = p.T;
        vxb = p.u;
        vyb = p.v;
        vzb = p.w;
        return;
    }
    const double R_out = un_in + 2.0 * a_in * inv_gm1;
    const double R_in = un_inf - 2.0 * a_inf * inv_gm1;
    const double un_b = 0.5 * (R_out + R_in);
    double a_b = 0.25 * gm1 * (R_out - R_in);
    if (a_b <= 0.0) a_b = a_inf;
    double s = 0.0;
    if (un_b >= 0.0) {
        vxb = vx + (un_b - un_in) * nx;
        vyb = vy + (un_b - un_in) * ny;
        vzb = vz + (un_b - un_in) * nz;
        const double rho_in = p_in / (R * std::max(T_in, 1.0e-6));
        s = p_in / std::pow(std::max(rho_in, 1.0e-12), gamma);
    } else {
        vxb = p.u + (un_b - un_inf) * nx;
        vyb = p.v + (un_b - un_inf) * ny;
        vzb = p.w + (un_b - un_inf) * nz;
        const double rho_inf = p.p / (R * std::max(p.T, 1.0e-6));
        s = p.p / std::pow(std::max(rho_inf, 1.0e-12), gamma);
    }
    const double rho_b = std::pow((a_b * a_b) / (gamma * std::max(s, 1.0e-12)), inv_gm1);
    pb = (rho_b * a_b * a_b) / gamma;
    Tb = (a_b * a_b) / (gamma * R);
}

// boundary_gradients.hpp:15-30, 71-86, 118-131
void grad_dirichlet(double* g, const double* gin, double qin, double qfix, double nx, double ny, double nz, double rinv) {
    const double ddn = (qfix - qin) * rinv;
    const double gn = gin[0] * nx + gin[1] * ny + gin[2] * nz;
    const double fac = 2.0 * (ddn - gn);
    g[0] = gin[0] + fac * nx;
    g[1] = gin[1] + fac * ny;
    g[2] = gin[2] + fac * nz;
}
void grad_neumann(double* g, const double* gin, double fixed, double nx, double ny, double nz) {
    const double gn = gin[0] * nx + gin[1] * ny + gin[2] * nz;
    const double fac = 2.0 * (fixed - gn);
    g[0] = gin[0] + fac * nx;
    g[1] = gin[1] + fac * ny;
    g[2] = gin[2] + fac * nz;
}
void grad_slip(double* g, const double* gin, double qin, double qgh, double nx, double ny, double nz, double rinv) {
    const double gn = gin[0] * nx + gin[1] * ny + gin[2] * nz;
    const double ddn = 0.5 * (qgh - qin) * rinv;
    const double fac = 2.0 * (ddn - gn);
    g[0] = gin[0] + fac * nx;
    g[1] = gin[1] + fac * ny;
    g[2] = gin[2] + fac * nz;
}

// bc/physical/*.hpp value kernels, dispatched per patch (bc_manager.hpp:185-191)
void update_ghosts(oracle_ctx& c) {
    double* Q[5];
    for (int v = 0; v < 5; ++v) Q[v] = plane(c.q, c.ntot, v);
    for (const BcParams& bp : c.bcs) {
        for (i32 f = bp.fbeg; f < bp.fend; ++f) {
            const size_t fs = static_cast<size_t>(f);
            const size_t in = static_cast<size_t>(c.owner[fs]);
            const size_t gh = static_cast<size_t>(c.nc) + (fs - static_cast<size_t>(c.ninner));
            const double nx = c.nx[fs], ny = c.ny[fs], nz = c.nz[fs];
            switch (bp.type) {
                case SupIn: {  // supersonic_inlet.hpp:38-64
                    const double qi[5] = {bp.p, bp.u, bp.v, bp.w, bp.T};
                    for (int v = 0; v < 5; ++v) Q[v][gh] = 2.0 * qi[v] - Q[v][in];
                    Q[0][gh] = std::max(Q[0][gh], 1.0);
                    Q[4][gh] = std::max(Q[4][gh], 1.0);
                    break;
                }
                case SupOut:  // supersonic_outlet.hpp:25-45
                    for (int v = 0; v < 5; ++v) Q[v][gh] = Q[v][in];
                    break;
                case Slip:
                case Sym: {  // slip_wall.hpp:25-52, symmetry.hpp:25-52
                    const double vn2 = 2.0 * (Q[1][in] * nx + Q[2][in] * ny + Q[3][in] * nz);
                    Q[1][gh] = Q[1][in] - vn2 * nx;
                    Q[2][gh] = Q[2][in] - vn2 * ny;
                    Q[3][gh] = Q[3][in] - vn2 * nz;
                    Q[0][gh] = Q[0][in];
                    Q[4][gh] = Q[4][in];
                    break;
                }
                case Far: {  // farfield.hpp:120-165
                    double pb = 0, Tb = 0, vxb = 0, vyb = 0, vzb = 0;
                    bool so = false;
                    farfield_state(bp, c.eos.g, c.eos.R, Q[0][in], Q[4][in], Q[1][in], Q[2][in], Q[3][in], nx, ny, nz, pb, Tb,
                                   vxb, vyb, vzb, so);
                    if (so) {
                        for (int v = 0; v < 5; ++v) Q[v][gh] = Q[v][in];
                    } else {
                        const double qb[5] = {pb, vxb, vyb, vzb, Tb};
                        for (int v = 0; v < 5; ++v) Q[v][gh] = 2.0 * qb[v] - Q[v][in];
                        Q[0][gh] = std::max(Q[0][gh], 1.0);
                        Q[4][gh] = std::max(Q[4][gh], 1.0);
                    }
                    break;
                }
                case NoSlip: {  // noslip_wall.hpp:38-75
                    Q[0][gh] = Q[0][in];
                    Q[1][gh] = 2.0 * bp.u - Q[1][in];
                    Q[2][gh] = 2.0 * bp.v - Q[2][in];
                    Q[3][gh] = 2.0 * bp.w - Q[3][in];
                    Q[4][gh] = 2.0 * bp.T - Q[4][in];
                    Q[4][gh] = std::max(Q[4][gh], 1.0);
                    break;
                }
                case NoSlipHF: {  // noslip_wall_heat_flux.hpp:56-110 (compressible branch)
                    const double rcfn = (c.fx[fs] - c.cx[in]) * nx + (c.fy[fs] - c.cy[in]) * ny + (c.fz[fs] - c.cz[in]) * nz;
                    const double rinv = 1.0 / std::max(rcfn, 1.0e-14);
                    Q[0][gh] = Q[0][in];
                    Q[1][gh] = 2.0 * bp.u - Q[1][in];
                    Q[2][gh] = 2.0 * bp.v - Q[2][in];
                    Q[3][gh] = 2.0 * bp.w - Q[3][in];
                    const double r = 1.0 / rinv;
                    Q[4][gh] = Q[4][in] + 2.0 * bp.tmp_grad * r;
                    Q[4][gh] = std::max(Q[4][gh], 1.0);
                    break;
                }
                case SubIn: {  // subsonic_inlet.hpp:38-75
                    Q[0][gh] = Q[0][in];
                    Q[1][gh] = 2.0 * bp.u - Q[1][in];
                    Q[2][gh] = 2.0 * bp.v - Q[2][in];
                    Q[3][gh] = 2.0 * bp.w - Q[3][in];
                    Q[4][gh] = 2.0 * bp.T - Q[4][in];
                    Q[4][gh] = std::max(Q[4][gh], 1.0);
                    break;
                }
                case SubOut: {  // subsonic_outlet.hpp:34-60
                    Q[0][gh] = 2.0 * bp.p - Q[0][in];
                    Q[0][gh] = std::max(Q[0][gh], 1.0);
                    for (int v = 1; v < 5; ++v) Q[v][gh] = Q[v][in];
                    break;
                }
                default:
                    break;
            }
        }
    }
}

// bc/physical/*.hpp gradient kernels (bc_manager.hpp:201-210)
void update_ghost_grads(oracle_ctx& c) {
    double *GX[5], *GY[5], *GZ[5], *Q[5];
    for (int v = 0; v < 5; ++v) {
        GX[v] = plane(c.gx, c.ntot, v);
        GY[v] = plane(c.gy, c.ntot, v);
        GZ[v] = plane(c.gz, c.ntot, v);
        Q[v] = plane(c.q, c.ntot, v);
    }
    for (const BcParams& bp : c.bcs) {
        for (i32 f = bp.fbeg; f < bp.fend; ++f) {
            const size_t fs = static_cast<size_t>(f);
            const size_t in = static_cast<size_t>(c.owner[fs]);
            const size_t gh = static_cast<size_t>(c.nc) + (fs - static_cast<size_t>(c.ninner));
            const double nx = c.nx[fs], ny = c.ny[fs], nz = c.nz[fs];
            // boundary_traversal.hpp:76-119
            const double rcfn = (c.fx[fs] - c.cx[in]) * nx + (c.fy[fs] - c.cy[in]) * ny + (c.fz[fs] - c.cz[in]) * nz;
            const double rinv = 1.0 / std::max(rcfn, 1.0e-14);
            double gin[5][3], gout[5][3];
            for (int v = 0; v < 5; ++v) {
                gin[v][0] = GX[v][in];
                gin[v][1] = GY[v][in];
                gin[v][2] = GZ[v][in];
                gout[v][0] = gin[v][0];
                gout[v][1] = gin[v][1];
                gout[v][2] = gin[v][2];
            }
            switch (bp.type) {
                case SupIn: {
                    const double qi[5] = {bp.p, bp.u, bp.v, bp.w, bp.T};
                    for (int v = 0; v < 5; ++v) grad_dirichlet(gout[v], gin[v], Q[v][in], qi[v], nx, ny, nz, rinv);
                    break;
                }
                case SupOut:
                    break;
                case Slip:
                case Sym:
                    grad_neumann(gout[0], gin[0], 0.0, nx, ny, nz);
                    for (int v = 1; v < 4; ++v) grad_slip(gout[v], gin[v], Q[v][in], Q[v][gh], nx, ny, nz, rinv);
                    grad_neumann(gout[4], gin[4], 0.0, nx, ny, nz);
                    break;
                case Far: {
                    double pb = 0, Tb = 0, vxb = 0, vyb = 0, vzb = 0;
                    bool so = false;
                    farfield_state(bp, c.eos.g, c.eos.R, Q[0][in], Q[4][in], Q[1][in], Q[2][in], Q[3][in], nx, ny, nz, pb, Tb,
                                   vxb, vyb, vzb, so);
                    if (!so) {
                        const double qb[5] = {pb, vxb, vyb, vzb, Tb};
                        for (int v = 0; v < 5; ++v) grad_dirichlet(gout[v], gin[v], Q[v][in], qb[v], nx, ny, nz, rinv);
                    }
                    break;
                }
                case NoSlip:
                case SubIn: {
                    const double qw[4] = {bp.u, bp.v, bp.w, bp.T};
                    for (int v = 1; v < 5; ++v) grad_dirichlet(gout[v], gin[v], Q[v][in], qw[v - 1], nx, ny, nz, rinv);
                    break;
                }
                case NoSlipHF: {
                    const double qw[3] = {bp.u, bp.v, bp.w};
                    for (int v = 1; v < 4; ++v) grad_dirichlet(gout[v], gin[v], Q[v][in], qw[v - 1], nx, ny, nz, rinv);
                    grad_neumann(gout[4], gin[4], bp.tmp_grad, nx, ny, nz);
                    break;
                }
                case SubOut:
                    grad_dirichlet(gout[0], gin[0], Q[0][in], bp.p, nx, ny, nz, rinv);
                    break;
                default:
                    break;
            }
            for (int v = 0; v < 5; ++v) {
                GX[v][gh] = gout[v][0];
                GY[v][gh] = gout[v][1];
                GZ[v][gh] = gout[v][2];
            }
        }
    }
}

// gradient_operator.cpp:197-278 (Green-Gauss face based)
void gradient_ggfb(oracle_ctx& c) {
    for (int v = 0; v < 5; ++v) {
        std::fill_n(plane(c.gx, c.ntot, v), c.nc, 0.0);
        std::fill_n(plane(c.gy, c.ntot, v), c.nc, 0.0);
        std::fill_n(plane(c.gz, c.ntot, v), c.nc, 0.0);
    }
    for (i32 f = 0; f < c.nf; ++f) {
        const size_t fs = static_cast<size_t>(f);
        const size_t o = static_cast<size_t>(c.owner[fs]);
        const bool inner = f < c.ninner;
        const size_t n = inner ? static_cast<size_t>(c.neigh[fs]) : static_cast<size_t>(c.nc) + fs - static_cast<size_t>(c.ninner);
        const double a05 = 0.5 * c.area[fs];
        const double hx = a05 * c.nx[fs], hy = a05 * c.ny[fs], hz = a05 * c.nz[fs];
        for (int v = 0; v < 5; ++v) {
            const double* s = plane(c.q, c.ntot, v);
            const double sum = s[o] + s[n];
            const double fx = hx * sum, fy = hy * sum, fz = hz * sum;
            plane(c.gx, c.ntot, v)[o] += fx;
            plane(c.gy, c.ntot, v)[o] += fy;
            plane(c.gz, c.ntot, v)[o] += fz;
            if (inner) {
                plane(c.gx, c.ntot, v)[n] -= fx;
                plane(c.gy, c.ntot, v)[n] -= fy;
                plane(c.gz, c.ntot, v)[n] -= fz;
            }
        }
    }
    for (i32 i = 0; i < c.nc; ++i) {
        const double iv = 1.0 / c.vol[static_cast<size_t>(i)];
        for (int v = 0; v < 5; ++v) {
            plane(c.gx, c.ntot, v)[i] *= iv;
            plane(c.gy, c.ntot, v)[i] *= iv;
            plane(c.gz, c.ntot, v)[i] *= iv;
        }
    }
}

// gradient_operator.cpp:61-130 (Green-Gauss cell based)
void gradient_ggcb(oracle_ctx& c) {
    for (i32 i = 0; i < c.nc; ++i) {
        double g[3][5] = {};
        for (i32 e = c.cf_off[static_cast<size_t>(i)]; e < c.cf_off[static_cast<size_t>(i) + 1]; ++e) {
            const size_t f = static_cast<size_t>(c.cf[static_cast<size_t>(e)]);
            const i32 o = c.owner[f], nb = c.neigh[f];
            const double s05 = (i == o) ? 0.5 : -0.5;
            const double gxf = s05 * c.area[f] * c.nx[f];
            const double gyf = s05 * c.area[f] * c.ny[f];
            const double gzf = s05 * c.area[f] * c.nz[f];
            const size_t j = (nb >= 0) ? (i == o ? static_cast<size_t>(nb) : static_cast<size_t>(o))
                                       : static_cast<size_t>(c.nc) + f - static_cast<size_t>(c.ninner);
            for (int v = 0; v < 5; ++v) {
                const double* s = plane(c.q, c.ntot, v);
                const double sf = s[static_cast<size_t>(i)] + s[j];
                g[0][v] += gxf * sf;
                g[1][v] += gyf * sf;
                g[2][v] += gzf * sf;
            }
        }
        const double iv = 1.0 / c.vol[static_cast<size_t>(i)];
        for (int v = 0; v < 5; ++v) {
            plane(c.gx, c.ntot, v)[i] = g[0][v] * iv;
            plane(c.gy, c.ntot, v)[i] = g[1][v] * iv;
            plane(c.gz, c.ntot, v)[i] = g[2][v] * iv;
        }
    }
}

// gradient_operator.cpp:549-599 (LSQ node stencil apply)
void gradient_lsqn(oracle_ctx& c) {
    const size_t E = c.ccn.size();
    const double* wx = c.lsq_w.data();
    const double* wy = wx + E;
    const double* wz = wy + E;
    for (i32 i = 0; i < c.nc; ++i) {
        double g[3][5] = {};
        double self[5];
        for (int v = 0; v < 5; ++v) self[v] = plane(c.q, c.ntot, v)[i];
        for (i32 j = c.ccn_off[static_cast<size_t>(i)]; j < c.ccn_off[static_cast<size_t>(i) + 1]; ++j) {
            const size_t nb = static_cast<size_t>(c.ccn[static_cast<size_t>(j)]);
            for (int v = 0; v < 5; ++v) {
                const double d = plane(c.q, c.ntot, v)[nb] - self[v];
                g[0][v] += wx[j] * d;
                g[1][v] += wy[j] * d;
                g[2][v] += wz[j] * d;
            }
        }
        for (int v = 0; v < 5; ++v) {
            plane(c.gx, c.ntot, v)[i] = g[0][v];
            plane(c.gy, c.ntot, v)[i] = g[1][v];
            plane(c.gz, c.ntot, v)[i] = g[2][v];
        }
    }
}

// muscl.hpp:26-104
void compute_limiters(oracle_ctx& c) {
    const double k3 = c.num.venkat_k * c.num.venkat_k * c.num.venkat_k;
    for (i32 i = 0; i < c.nc; ++i) {
        const size_t is = static_cast<size_t>(i);
        double qmax[5], qmin[5], qc[5];
        for (int v = 0; v < 5; ++v) {
            qc[v] = plane(c.q, c.ntot, v)[is];
            qmax[v] = qc[v];
            qmin[v] = qc[v];
        }
        for (i32 j = c.ccf_off[is]; j < c.ccf_off[is + 1]; ++j) {
            const size_t js = static_cast<size_t>(c.ccf[static_cast<size_t>(j)]);
            for (int v = 0; v < 5; ++v) {
                const double qn = plane(c.q, c.ntot, v)[js];
                qmax[v] = std::max(qmax[v], qn);
                qmin[v] = std::min(qmin[v], qn);
            }
        }
        double ph[5] = {1.0, 1.0, 1.0, 1.0, 1.0};
        const double eps2 = k3 * c.vol[is];
        for (i32 e = c.cf_off[is]; e < c.cf_off[is + 1]; ++e) {
            const size_t f = static_cast<size_t>(c.cf[static_cast<size_t>(e)]);
            const double ex = c.fx[f] - c.cx[is], ey = c.fy[f] - c.cy[is], ez = c.fz[f] - c.cz[is];
            for (int v = 0; v < 5; ++v) {
                const double df = plane(c.gx, c.ntot, v)[is] * ex + plane(c.gy, c.ntot, v)[is] * ey + plane(c.gz, c.ntot, v)[is] * ez;
                const double dnb = (df > 0.0) ? (qmax[v] - qc[v]) : (qmin[v] - qc[v]);
                ph[v] = std::min(ph[v], limiter_phi(c.num.limiter, dnb, df, eps2));
            }
        }
        for (int v = 0; v < 5; ++v) plane(c.phi, c.ntot, v)[is] = ph[v];
    }
}

// residual_kernel.hpp:64-91, 175-306
void compute_fluxes(oracle_ctx& c) {
    for (int v = 0; v < 5; ++v) std::fill_n(plane(c.res, c.ntot, v), c.ntot, 0.0);
    std::fill(c.lam.begin(), c.lam.end(), 0.0);
    std::fill(c.fscale.begin(), c.fscale.end(), 0.0);
    const bool muscl = c.num.reconstruction == 1;
    const bool visc = c.num.flow_model == 1;
    Views vv;
    for (int v = 0; v < 5; ++v) {
        vv.q[v] = plane(c.q, c.ntot, v);
        vv.gx[v] = plane(c.gx, c.ntot, v);
        vv.gy[v] = plane(c.gy, c.ntot, v);
        vv.gz[v] = plane(c.gz, c.ntot, v);
    }
    for (i32 f = 0; f < c.nf; ++f) {
        const size_t fs = static_cast<size_t>(f);
        const bool inner = f < c.ninner;
        const size_t c0 = static_cast<size_t>(c.owner[fs]);
        const size_t c1 = inner ? static_cast<size_t>(c.neigh[fs]) : static_cast<size_t>(c.nc) + fs - static_cast<size_t>(c.ninner);
        const double nxf = c.nx[fs], nyf = c.ny[fs], nzf = c.nz[fs], Af = c.area[fs];
        double qL[5], qR[5];
        if (inner && muscl) {  // muscl.hpp:108-141
            const double d0x = c.fx[fs] - c.cx[c0], d0y = c.fy[fs] - c.cy[c0], d0z = c.fz[fs] - c.cz[c0];
            const double d1x = c.fx[fs] - c.cx[c1], d1y = c.fy[fs] - c.cy[c1], d1z = c.fz[fs] - c.cz[c1];
            for (int v = 0; v < 5; ++v) {
                const double dot0 = vv.gx[v][c0] * d0x + vv.gy[v][c0] * d0y + vv.gz[v][c0] * d0z;
                qL[v] = vv.q[v][c0] + plane(c.phi, c.ntot, v)[c0] * dot0;
                const double dot1 = vv.gx[v][c1] * d1x + vv.gy[v][c1] * d1y + vv.gz[v][c1] * d1z;
                qR[v] = vv.q[v][c1] + plane(c.phi, c.ntot, v)[c1] * dot1;
            }
        } else {  // first order / boundary faces (muscl.hpp:143-157)
            for (int v = 0; v < 5; ++v) {
                qL[v] = vv.q[v][c0];
                qR[v] = vv.q[v][c1];
            }
        }
        double UL[5], UR[5];
        prim_to_cons(c.eos, qL[0], qL[1], qL[2], qL[3], qL[4], UL);
        prim_to_cons(c.eos, qR[0], qR[1], qR[2], qR[3], qR[4], UR);
        double F[5];
        double smax = 0.0;
        hllc(c.eos, UL, UR, nxf, nyf, nzf, Af, F, smax);
        if (c.sa_on()) c.sa.mdot[fs] = F[0];  // residual_kernel.hpp:222-224, 275-277
        double lam_f = Af * smax;
        if (visc) {
            double Fv[5] = {0, 0, 0, 0, 0};
            double lv = 0.0;
            // eddy viscosity of the two sides, zero in BC ghosts (residual_kernel.hpp:229-230, 282-283)
            const double mutL = c.mut_at(c0);
            const double mutR = c.mut_at(c1);
            viscous_flux(c.eos, c.dinv[fs], c.dvx[fs], c.dvy[fs], c.dvz[fs], c0, c1, vv, mutL, mutR, nxf, nyf, nzf, Af, Fv, lv,
                         c.num.prandtl, c.num.prandtl_turb);
            for (int v = 1; v < 5; ++v) F[v] += Fv[v];
            lam_f += lv;
        }
        for (int v = 0; v < 5; ++v) plane(c.res, c.ntot, v)[c0] += F[v];
        if (!inner)
            for (int v = 0; v < 5; ++v) c.bflux[(fs - static_cast<size_t>(c.ninner)) * 5 + static_cast<size_t>(v)] = F[v];
        c.lam[c0] += lam_f;
        for (int v = 0; v < 5; ++v) plane(c.fscale, c.ntot, v)[c0] += std::fabs(F[v]);
        if (inner) {
            for (int v = 0; v < 5; ++v) plane(c.res, c.ntot, v)[c1] -= F[v];
            c.lam[c1] += lam_f;
            for (int v = 0; v < 5; ++v) plane(c.fscale, c.ntot, v)[c1] += std::fabs(F[v]);
        }
    }
}

// --- Spalart-Allmaras module (spalart_allmaras.hpp) --------------------------
// The reference header does not compile (it includes headers that do not exist
// and is never dispatched), so this restatement follows its equations and
// call order with three corrections, each marked "correction":
//   1. the diffusive face flux leaves the owner as -coef (grad nu . n) A
//      (the header adds +coef (...) A, an anti-diffusion);
//   2. the cb2 term is a source, res -= V cb2/sigma rho |grad nu|^2, as the
//      header's own comment states (its code adds it with the sign of D);
//   3. the "mass flux < 0 -> freestream" override of a boundary face applies
//      to inflow / outflow zones only; at walls and symmetry planes the zone
//      ghost is used whatever the sign of a round-off mass flux;
//   4. fw = g ((1 + cw3^6) / (g^6 + cw3^6))^(1/6) as the header's own formula
//      (spalart_allmaras.hpp:17) states; its code drops the cw3^6 of the
//      denominator (spalart_allmaras.hpp:419), which gives fw(r = 1) = 2.005.
// BC ghosts of nu_tilde carry rnu_g / rho_g for the Green-Gauss gradients
// (the header leaves them zero; the node LSQ stencil does not read them).
constexpr double kSaCb1 = 0.1355, kSaCb2 = 0.622, kSaSigma = 2.0 / 3.0, kSaKappa = 0.41;
constexpr double kSaCw2 = 0.3, kSaCw3 = 2.0, kSaCv1 = 7.1;
constexpr double kSaCw1 = kSaCb1 / (kSaKappa * kSaKappa) + (1.0 + kSaCb2) / kSaSigma;
constexpr double kSaCv1_3 = kSaCv1 * kSaCv1 * kSaCv1;
constexpr double kSaSbarFloor = 1.0e-10, kSaMinWallDist = 1.0e-10;
enum NuZone { ZWall = 0, ZSymmetry, ZInflow, ZOutflow };
void wall_distance(oracle_ctx& c, int32_t stencil, int32_t max_sweeps, double rel_tol, double* dist);

int nu_zone(int bc_type) {  // spalart_allmaras.hpp:447-464
    switch (bc_type) {
        case NoSlip:
        case NoSlipHF: return ZWall;
        case SupIn:
        case SubIn:
        case Far: return ZInflow;
        case SupOut:
        case SubOut: return ZOutflow;
        default: return ZSymmetry;
    }
}

double sa_fv1(double chi) {
    const double chi3 = chi * chi * chi;
    return chi3 / (chi3 + kSaCv1_3);
}

// cell_sources (spalart_allmaras.hpp:382-432) for one cell: P, D, the cb2
// term and the pieces of the point-implicit diagonal
struct SaSource {
    double P, D, cb2, fv1, fv2, sbar, r, fw;
};
SaSource sa_source(double rho, double nu, double mu_lam, double d_in, double S, double grad2) {
    SaSource o{};
    const double inv_k2 = 1.0 / (kSaKappa * kSaKappa);
    const double d = std::max(d_in, kSaMinWallDist);
    const double inv_d2 = 1.0 / (d * d);
    const double nu_lam = mu_lam / rho;
    const double chi = nu / nu_lam;
    o.fv1 = sa_fv1(chi);
    o.fv2 = 1.0 - chi / (1.0 + chi * o.fv1);
    o.sbar = std::max(S + nu * o.fv2 * inv_k2 * inv_d2, kSaSbarFloor);
    double r = nu / (o.sbar * kSaKappa * kSaKappa * d * d);
    r = std::min(std::max(r, 0.0), 10.0);
    o.r = r;
    const double r2 = r * r;
    const double r6 = r2 * r2 * r2;
    const double g = r + kSaCw2 * (r6 - r);
    const double g2 = std::max(g * g, kSaSbarFloor);
    const double g6 = g2 * g2 * g2;
    const double cw3_6 = kSaCw3 * kSaCw3 * kSaCw3 * kSaCw3 * kSaCw3 * kSaCw3;
    o.fw = g * std::sqrt(std::cbrt((1.0 + cw3_6) / (g6 + cw3_6)));  // correction 4
    o.P = kSaCb1 * o.sbar * rho * nu;
    o.D = kSaCw1 * o.fw * rho * nu * nu * inv_d2;
    o.cb2 = (kSaCb2 / kSaSigma) * rho * grad2;
    return o;
}

// Strain-rate magnitude sqrt(2 S_ij S_ij) from the velocity gradient rows
double strain_rate(const oracle_ctx& c, size_t i) {
    const size_t n = static_cast<size_t>(c.ntot);
    const auto G = [&](const std::vector<double>& a, int v) { return a[static_cast<size_t>(v) * n + i]; };
    const double ux = G(c.gx, 1), uy = G(c.gy, 1), uz = G(c.gz, 1);
    const double vx = G(c.gx, 2), vy = G(c.gy, 2), vz = G(c.gz, 2);
    const double wx = G(c.gx, 3), wy = G(c.gy, 3), wz = G(c.gz, 3);
    const double s_xy = uy + vx, s_xz = uz + wx, s_yz = vz + wy;
    const double s2 = 2.0 * (ux * ux + vy * vy + wz * wz) + 0.5 * (s_xy * s_xy + s_xz * s_xz + s_yz * s_yz);
    return std::sqrt(std::max(s2, 0.0));
}

// apply_bcs (spalart_allmaras.hpp:215-240): ghost rho nu_tilde per zone
void sa_apply_bcs(oracle_ctx& c) {
    const double* Q0 = plane(c.q, c.ntot, 0);
    const double* Q4 = plane(c.q, c.ntot, 4);
    for (i32 f = c.ninner; f < c.nf; ++f) {
        const size_t i = static_cast<size_t>(f - c.ninner);
        const size_t c0 = static_cast<size_t>(c.owner[static_cast<size_t>(f)]);
        const size_t cg = static_cast<size_t>(c.nc) + i;
        switch (c.sa.kind[i]) {
            case ZWall: c.sa.rnu[cg] = -c.sa.rnu[c0]; break;
            case ZInflow: c.sa.rnu[cg] = c.eos.density_Tp(Q4[cg], Q0[cg]) * c.sa.nu_inf; break;
            default: c.sa.rnu[cg] = c.sa.rnu[c0]; break;
        }
    }
}

// pre_sweep (spalart_allmaras.hpp:246-270): density, nu_tilde, eddy viscosity
void sa_pre_sweep(oracle_ctx& c) {
    const double* Q0 = plane(c.q, c.ntot, 0);
    const double* Q4 = plane(c.q, c.ntot, 4);
    for (i32 i = 0; i < c.nc; ++i) {
        const size_t s = static_cast<size_t>(i);
        const double rho = c.eos.density_Tp(Q4[s], Q0[s]);
        const double nu = c.sa.rnu[s] / rho;
        c.sa.rho[s] = rho;
        c.sa.nu[s] = nu;
        const double nu_lam = viscosity(Q4[s]) / rho;
        c.sa.mut[s] = rho * nu * sa_fv1(nu / nu_lam);
    }
    for (i32 g = c.nc; g < c.ntot; ++g) {
        const size_t s = static_cast<size_t>(g);
        const double rho_g = c.eos.density_Tp(Q4[s], Q0[s]);
        c.sa.rho[s] = rho_g;
        c.sa.nu[s] = c.sa.rnu[s] / rho_g;
        c.sa.mut[s] = 0.0;
    }
}

// compute_gradients (spalart_allmaras.hpp:196-212): the configured gradient
// method on nu_tilde, then BC ghost planes mirror the owner (sign flip at walls)
void sa_gradients(oracle_ctx& c) {
    const size_t n = static_cast<size_t>(c.ntot);
    const double* s = c.sa.nu.data();
    double* gx = c.sa.g.data();
    double* gy = gx + n;
    double* gz = gy + n;
    if (c.num.gradient == 2) {  // gradient_operator.cpp:549-599
        const size_t E = c.ccn.size();
        const double* wx = c.lsq_w.data();
        const double* wy = wx + E;
        const double* wz = wy + E;
        for (i32 i = 0; i < c.nc; ++i) {
            const size_t is = static_cast<size_t>(i);
            double a = 0, b = 0, d = 0;
            for (i32 j = c.ccn_off[is]; j < c.ccn_off[is + 1]; ++j) {
                const double dv = s[static_cast<size_t>(c.ccn[static_cast<size_t>(j)])] - s[is];
                a += wx[j] * dv;
                b += wy[j] * dv;
                d += wz[j] * dv;
            }
            gx[is] = a;
            gy[is] = b;
            gz[is] = d;
        }
    } else if (c.num.gradient == 1) {  // gradient_operator.cpp:61-130
        for (i32 i = 0; i < c.nc; ++i) {
            const size_t is = static_cast<size_t>(i);
            double a = 0, b = 0, d = 0;
            for (i32 e = c.cf_off[is]; e < c.cf_off[is + 1]; ++e) {
                const size_t f = static_cast<size_t>(c.cf[static_cast<size_t>(e)]);
                const i32 o = c.owner[f], nb = c.neigh[f];
                const double s05 = (i == o) ? 0.5 : -0.5;
                const size_t j = (nb >= 0) ? (i == o ? static_cast<size_t>(nb) : static_cast<size_t>(o))
                                           : static_cast<size_t>(c.nc) + f - static_cast<size_t>(c.ninner);
                const double sf = s[is] + s[j];
                a += s05 * c.area[f] * c.nx[f] * sf;
                b += s05 * c.area[f] * c.ny[f] * sf;
                d += s05 * c.area[f] * c.nz[f] * sf;
            }
            const double iv = 1.0 / c.vol[is];
            gx[is] = a * iv;
            gy[is] = b * iv;
            gz[is] = d * iv;
        }
    } else {  // gradient_operator.cpp:11-59
        std::fill_n(gx, c.nc, 0.0);
        std::fill_n(gy, c.nc, 0.0);
        std::fill_n(gz, c.nc, 0.0);
        for (i32 f = 0; f < c.nf; ++f) {
            const size_t fs = static_cast<size_t>(f);
            const size_t o = static_cast<size_t>(c.owner[fs]);
            const bool inner = f < c.ninner;
            const size_t nb = inner ? static_cast<size_t>(c.neigh[fs]) : static_cast<size_t>(c.nc) + fs - static_cast<size_t>(c.ninner);
            const double a05 = 0.5 * c.area[fs];
            const double sum = s[o] + s[nb];
            const double fx = a05 * c.nx[fs] * sum, fy = a05 * c.ny[fs] * sum, fz = a05 * c.nz[fs] * sum;
            gx[o] += fx;
            gy[o] += fy;
            gz[o] += fz;
            if (inner) {
                gx[nb] -= fx;
                gy[nb] -= fy;
                gz[nb] -= fz;
            }
        }
        for (i32 i = 0; i < c.nc; ++i) {
            const double iv = 1.0 / c.vol[static_cast<size_t>(i)];
            gx[i] *= iv;
            gy[i] *= iv;
            gz[i] *= iv;
        }
    }
    for (i32 f = c.ninner; f < c.nf; ++f) {
        const size_t i = static_cast<size_t>(f - c.ninner);
        const size_t c0 = static_cast<size_t>(c.owner[static_cast<size_t>(f)]);
        const size_t cg = static_cast<size_t>(c.nc) + i;
        const double sg = (c.sa.kind[i] == ZWall) ? -1.0 : 1.0;
        gx[cg] = sg * gx[c0];
        gy[cg] = sg * gy[c0];
        gz[cg] = sg * gz[c0];
    }
}

// face_gradient (spalart_allmaras.hpp:500-525): averaged + skew-corrected
void sa_face_gradient(const oracle_ctx& c, size_t f, size_t cl, double nu_l, size_t cr, double nu_r, double gn[3]) {
    const size_t n = static_cast<size_t>(c.ntot);
    const double* g = c.sa.g.data();
    const double dx = c.dvx[f], dy = c.dvy[f], dz = c.dvz[f];
    gn[0] = 0.5 * (g[cl] + g[cr]);
    gn[1] = 0.5 * (g[n + cl] + g[n + cr]);
    gn[2] = 0.5 * (g[2 * n + cl] + g[2 * n + cr]);
    const double jump = nu_r - nu_l;
    const double gd = gn[0] * dx + gn[1] * dy + gn[2] * dz;
    const double inv_d = c.dinv[f];
    const double corr = (jump - gd) * inv_d * inv_d;
    gn[0] += corr * dx;
    gn[1] += corr * dy;
    gn[2] += corr * dz;
}

// face_sweep (spalart_allmaras.hpp:279-373): upwind convection on the mean
// mass flux + diffusion; interior faces add their spectral radius to lam.
// diag collects the point-implicit diagonal: outflow m / rho and coef A / (d rho)
// of each face.
void sa_face_sweep(oracle_ctx& c) {
    std::fill(c.sa.res.begin(), c.sa.res.end(), 0.0);
    std::fill(c.sa.diag.begin(), c.sa.diag.end(), 0.0);
    const double inv_sigma = 1.0 / kSaSigma;
    const double* Q4 = plane(c.q, c.ntot, 4);
    const std::vector<double>& nu = c.sa.nu;
    const std::vector<double>& rho = c.sa.rho;
    for (i32 f = 0; f < c.ninner; ++f) {
        const size_t fs = static_cast<size_t>(f);
        const size_t c0 = static_cast<size_t>(c.owner[fs]), c1 = static_cast<size_t>(c.neigh[fs]);
        const double m = c.sa.mdot[fs];
        const double nu_up = (m >= 0.0) ? nu[c0] : nu[c1];
        const double f_conv = m * nu_up;
        double gn[3];
        sa_face_gradient(c, fs, c0, nu[c0], c1, nu[c1], gn);
        const double Tf = 0.5 * (Q4[c0] + Q4[c1]);
        const double mu_f = viscosity(Tf);
        const double rho_f = 0.5 * (rho[c0] + rho[c1]);
        const double nu_f = 0.5 * (nu[c0] + nu[c1]);
        const double coef = inv_sigma * (mu_f + rho_f * nu_f);
        const double f_diff = coef * (gn[0] * c.nx[fs] + gn[1] * c.ny[fs] + gn[2] * c.nz[fs]) * c.area[fs];
        const double flux = f_conv - f_diff;  // correction 1
        c.sa.res[c0] += flux;
        c.sa.res[c1] -= flux;
        const double visc_term = 2.0 * inv_sigma * (mu_f + rho_f * nu_f) / rho_f * c.area[fs] * c.dinv[fs];
        const double w = std::fabs(m) / rho_f + visc_term;
        c.lam[c0] += w;
        c.lam[c1] += w;
        const double dcoef = coef * c.area[fs] * c.dinv[fs];
        c.sa.diag[c0] += (std::max(m, 0.0) + dcoef) / rho[c0];
        c.sa.diag[c1] += (std::max(-m, 0.0) + dcoef) / rho[c1];
    }
    for (i32 f = c.ninner; f < c.nf; ++f) {
        const size_t fs = static_cast<size_t>(f);
        const size_t i = fs - static_cast<size_t>(c.ninner);
        const size_t c0 = static_cast<size_t>(c.owner[fs]);
        const size_t cg = static_cast<size_t>(c.nc) + i;
        const double m = c.sa.mdot[fs];
        const int kind = c.sa.kind[i];
        const bool inflow = m < 0.0 && (kind == ZInflow || kind == ZOutflow);  // correction 3
        const double nu_up = inflow ? c.sa.nu_inf : nu[c0];
        const double f_conv = m * nu_up;
        const double rho_g = rho[cg];
        const double nu_g = inflow ? c.sa.nu_inf : c.sa.rnu[cg] / rho_g;
        double gn[3];
        sa_face_gradient(c, fs, c0, nu[c0], cg, nu_g, gn);
        const double Tf = 0.5 * (Q4[c0] + Q4[cg]);
        const double mu_f = viscosity(Tf);
        const double rho_f = 0.5 * (rho[c0] + rho_g);
        const double coef = inv_sigma * (mu_f + rho_f * 0.5 * (nu[c0] + nu_g));
        const double f_diff = coef * (gn[0] * c.nx[fs] + gn[1] * c.ny[fs] + gn[2] * c.nz[fs]) * c.area[fs];
        c.sa.res[c0] += f_conv - f_diff;  // correction 1
        c.sa.diag[c0] += (std::max(m, 0.0) + coef * c.area[fs] * c.dinv[fs]) / rho[c0];
    }
}

// cell_sources (spalart_allmaras.hpp:382-432); correction 2 on the cb2 sign.
// diag gains V max(2 cw1 fw nu / d^2 - cb1 Sbar, 0).
void sa_cell_sources(oracle_ctx& c) {
    const double* Q4 = plane(c.q, c.ntot, 4);
    const size_t n = static_cast<size_t>(c.ntot);
    for (i32 i = 0; i < c.nc; ++i) {
        const size_t s = static_cast<size_t>(i);
        const double* g = c.sa.g.data();
        const double grad2 = g[s] * g[s] + g[n + s] * g[n + s] + g[2 * n + s] * g[2 * n + s];
        const SaSource o = sa_source(c.sa.rho[s], c.sa.nu[s], viscosity(Q4[s]), c.sa.dist[s], strain_rate(c, s), grad2);
        c.sa.res[s] += c.vol[s] * (o.D - o.P - o.cb2);
        const double d = std::max(c.sa.dist[s], kSaMinWallDist);
        const double lin = 2.0 * kSaCw1 * o.fw * c.sa.nu[s] / (d * d) - kSaCb1 * o.sbar;
        c.sa.diag[s] += c.vol[s] * std::max(lin, 0.0);
    }
}

// Module update after the mean-flow step.  Frozen: the reference's implicit
// Euler copies the module block unchanged (implicit_euler.hpp:64-72); point
// implicit: rnu -= res / (V / dt + diag).  Both end with the positivity clamp
// of post_stage (spalart_allmaras.hpp:435-442).
void sa_update(oracle_ctx& c) {
    for (i32 i = 0; i < c.nc; ++i) {
        const size_t s = static_cast<size_t>(i);
        double v = c.sa.rnu[s];
        if (c.sa.cfg.update == 1) v = v - c.sa.res[s] / (c.alpha[s] + c.sa.diag[s]);
        c.sa.rnu[s] = (v < 0.0) ? 0.0 : v;
    }
}

// solver.hpp:169-247
// BDF pseudo-source R += V (c0 U + c1 U^n [+ c2 U^{n-1}]) (solver.hpp:546-575)
void add_unsteady_source(oracle_ctx& c) {
    const double inv_dt = 1.0 / c.phys_dt;
    const size_t n = static_cast<size_t>(c.nc);
    for (int v = 0; v < 5; ++v) {
        const double* u = plane(c.U, c.ntot, v);
        const double* un = c.Un.data() + static_cast<size_t>(v) * n;
        const double* unm1 = c.Unm1.data() + static_cast<size_t>(v) * n;
        double* r = plane(c.res, c.ntot, v);
        if (c.bdf_cur == 2) {
            const double c0 = 1.5 * inv_dt, c1 = -2.0 * inv_dt, c2 = 0.5 * inv_dt;
            for (size_t i = 0; i < n; ++i) r[i] += c.vol[i] * (c0 * u[i] + c1 * un[i] + c2 * unm1[i]);
        } else {
            const double c0 = 1.0 * inv_dt, c1 = -1.0 * inv_dt;
            for (size_t i = 0; i < n; ++i) r[i] += c.vol[i] * (c0 * u[i] + c1 * un[i]);
        }
    }
}

void evaluate_residual(oracle_ctx& c) {
    for (i32 i = 0; i < c.nc; ++i) {
        const size_t is = static_cast<size_t>(i);
        const double Uc[5] = {plane(c.U, c.ntot, 0)[is], plane(c.U, c.ntot, 1)[is], plane(c.U, c.ntot, 2)[is],
                              plane(c.U, c.ntot, 3)[is], plane(c.U, c.ntot, 4)[is]};
        cons_to_prim(c.eos, Uc, plane(c.q, c.ntot, 0)[is], plane(c.q, c.ntot, 1)[is], plane(c.q, c.ntot, 2)[is],
                     plane(c.q, c.ntot, 3)[is], plane(c.q, c.ntot, 4)[is]);
    }
    update_ghosts(c);
    if (c.sa_on()) {
        sa_apply_bcs(c);
        sa_pre_sweep(c);
    }
    const bool need_grad = c.num.reconstruction == 1 || c.num.flow_model == 1;
    if (need_grad) {
        if (c.num.gradient == 0) gradient_ggfb(c);
        else if (c.num.gradient == 1) gradient_ggcb(c);
        else gradient_lsqn(c);
        update_ghost_grads(c);
        if (c.num.reconstruction == 1) compute_limiters(c);
        if (c.sa_on()) sa_gradients(c);
    }
    compute_fluxes(c);
    if (c.sa_on()) {
        sa_face_sweep(c);
        sa_cell_sources(c);
    }
    if (c.bdf_order > 0) add_unsteady_source(c);
    c.have_residual = true;
}

// solver.hpp:256-287, implicit branch; the unsteady diagonal term V c0 / dt
// with c0 = 1.5 (BDF2) or 1 (BDF1) in dual time stepping
void compute_dt(oracle_ctx& c) {
    const double inv_cfl = 1.0 / c.num.cfl;
    double unsteady = 0.0;
    if (c.bdf_order > 0) unsteady = ((c.bdf_cur == 2) ? 1.5 : 1.0) / c.phys_dt;
    for (i32 i = 0; i < c.nc; ++i) {
        const size_t is = static_cast<size_t>(i);
        const double l = std::max(c.lam[is], kLamFloor);
        c.dt[is] = c.num.cfl * c.vol[is] / l;
        c.alpha[is] = (l * inv_cfl) + (c.vol[is] * unsteady);
    }
}

// solver.hpp:256-287, explicit branch: alpha = dt / V = CFL / lambda
void compute_dt_explicit(oracle_ctx& c) {
    for (i32 i = 0; i < c.nc; ++i) {
        const size_t is = static_cast<size_t>(i);
        const double l = std::max(c.lam[is], kLamFloor);
        const double cfl_over_l = c.num.cfl / l;
        c.dt[is] = cfl_over_l * c.vol[is];
        c.alpha[is] = cfl_over_l;
    }
}

// jacobian_kernel.hpp:74-270
void assemble_jacobian(oracle_ctx& c) {
    std::fill(c.values.begin(), c.values.end(), 0.0);
    const bool visc = c.num.flow_model == 1;
    const double* Q[5];
    for (int v = 0; v < 5; ++v) Q[v] = plane(c.q, c.ntot, v);
    for (i32 f = 0; f < c.nf; ++f) {
        const size_t fs = static_cast<size_t>(f);
        const bool inner = f < c.ninner;
        const size_t c0 = static_cast<size_t>(c.owner[fs]);
        const size_t c1 = inner ? static_cast<size_t>(c.neigh[fs]) : static_cast<size_t>(c.nc) + fs - static_cast<size_t>(c.ninner);
        const double qL[5] = {Q[0][c0], Q[1][c0], Q[2][c0], Q[3][c0], Q[4][c0]};
        const double qR[5] = {Q[0][c1], Q[1][c1], Q[2][c1], Q[3][c1], Q[4][c1]};
        double UL[5], UR[5];
        prim_to_cons(c.eos, qL[0], qL[1], qL[2], qL[3], qL[4], UL);
        prim_to_cons(c.eos, qR[0], qR[1], qR[2], qR[3], qR[4], UR);
        double dFL[25], dFR[25];
        hllc_jacobian(c.eos, UL, UR, c.nx[fs], c.ny[fs], c.nz[fs], c.area[fs], dFL, dFR);
        if (visc) {
            double vL[25], vR[25];
            // jacobian_kernel.hpp:156-157, 209-210, 252-254: ghost side zero
            const double mutL = c.mut_at(c0);
            const double mutR = inner ? c.mut_at(c1) : 0.0;
            viscous_jacobian(c.eos, c.dinv[fs], qL, qR, mutL, mutR, c.area[fs], vL, vR, c.num.prandtl, c.num.prandtl_turb);
            for (int i = 0; i < 25; ++i) {
                dFL[i] += vL[i];
                if (inner) dFR[i] += vR[i];
            }
        }
        const size_t p00 = static_cast<size_t>(c.diag[c0]) * 25;
        if (inner) {
            const size_t p01 = bsr_pos(c, c0, static_cast<i32>(c1)) * 25;
            for (int i = 0; i < 25; ++i) {
                c.values[p00 + static_cast<size_t>(i)] += dFL[i];
                c.values[p01 + static_cast<size_t>(i)] += dFR[i];
            }
            const size_t p11 = static_cast<size_t>(c.diag[c1]) * 25;
            const size_t p10 = bsr_pos(c, c1, static_cast<i32>(c0)) * 25;
            for (int i = 0; i < 25; ++i) {
                c.values[p11 + static_cast<size_t>(i)] -= dFR[i];
                c.values[p10 + static_cast<size_t>(i)] -= dFL[i];
            }
        } else {
            for (int i = 0; i < 25; ++i) c.values[p00 + static_cast<size_t>(i)] += dFL[i];
        }
    }
    for (i32 i = 0; i < c.nc; ++i) {
        const size_t pos = static_cast<size_t>(c.diag[static_cast<size_t>(i)]) * 25;
        for (int k = 0; k < 5; ++k) c.values[pos + static_cast<size_t>(k) * 6] += c.alpha[static_cast<size_t>(i)];
    }
}

// bsr_helpers.hpp:139-203 (alpha = 1, beta = 0)
void spmv(const oracle_ctx& c, const double* x, double* y) {
    for (i32 i = 0; i < c.nc; ++i) {
        double acc[5] = {0, 0, 0, 0, 0};
        for (i32 k = c.row_ptr[static_cast<size_t>(i)]; k < c.row_ptr[static_cast<size_t>(i) + 1]; ++k) {
            const double* a = c.values.data() + static_cast<size_t>(k) * 25;
            const double* xb = x + static_cast<size_t>(c.cols[static_cast<size_t>(k)]) * 5;
            for (int ii = 0; ii < 5; ++ii) {
                double s = 0.0;
                for (int jj = 0; jj < 5; ++jj) s += a[ii * 5 + jj] * xb[jj];
                acc[ii] += s;
            }
        }
        for (int ii = 0; ii < 5; ++ii) y[static_cast<size_t>(i) * 5 + static_cast<size_t>(ii)] = acc[ii];
    }
}

struct Sgs {
    std::vector<double> lu;
    std::vector<i32> piv;
    std::vector<i32> colour;
    i32 n_colours = 0;
    int order = 1;
};

// preconditioners.cpp:62-93
bool sgs_setup(const oracle_ctx& c, Sgs& s) {
    const size_t n = static_cast<size_t>(c.nc);
    s.lu.assign(n * 25, 0.0);
    s.piv.assign(n * 5, 0);
    for (size_t i = 0; i < n; ++i) {
        std::copy_n(c.values.data() + static_cast<size_t>(c.diag[i]) * 25, 25, s.lu.data() + i * 25);
        if (!lu_factor5(s.lu.data() + i * 25, s.piv.data() + i * 5)) return false;
    }
    return true;
}

// One row of bsr_sgs_sweep<5> (bsr_helpers.hpp:390-435)
void sgs_row(const oracle_ctx& c, const Sgs& s, i32 i, const double* r, double* z) {
    const size_t is = static_cast<size_t>(i);
    const i32 d = c.diag[is];
    double acc[5];
    for (int ii = 0; ii < 5; ++ii) acc[ii] = r[is * 5 + static_cast<size_t>(ii)];
    for (i32 k = c.row_ptr[is]; k < c.row_ptr[is + 1]; ++k) {
        if (k == d) continue;
        const double* a = c.values.data() + static_cast<size_t>(k) * 25;
        const double* zb = z + static_cast<size_t>(c.cols[static_cast<size_t>(k)]) * 5;
        for (int ii = 0; ii < 5; ++ii) {
            double t = 0.0;
            for (int jj = 0; jj < 5; ++jj) t += a[ii * 5 + jj] * zb[jj];
            acc[ii] -= t;
        }
    }
    lu_solve5(s.lu.data() + is * 25, s.piv.data() + is * 5, acc);
    for (int ii = 0; ii < 5; ++ii) z[is * 5 + static_cast<size_t>(ii)] = acc[ii];
}

void colour_rows(const oracle_ctx& c, std::vector<i32>& colour, i32& n_colours) {
    colour.assign(static_cast<size_t>(c.nc), -1);
    n_colours = 0;
    std::vector<char> used;
    for (i32 i = 0; i < c.nc; ++i) {
        used.assign(static_cast<size_t>(n_colours) + 1, 0);
        for (i32 e = c.ccf_off[static_cast<size_t>(i)]; e < c.ccf_off[static_cast<size_t>(i) + 1]; ++e) {
            const i32 j = c.ccf[static_cast<size_t>(e)];
            if (j < i) used[static_cast<size_t>(colour[static_cast<size_t>(j)])] = 1;
        }
        i32 k = 0;
        while (used[static_cast<size_t>(k)]) ++k;
        colour[static_cast<size_t>(i)] = k;
        n_colours = std::max(n_colours, k + 1);
    }
}

// preconditioners.cpp:171-194 (z = 0, forward sweep, backward sweep)
void sgs_apply(const oracle_ctx& c, const Sgs& s, const double* r, double* z) {
    std::fill_n(z, static_cast<size_t>(c.nc) * 5, 0.0);
    if (s.order == 1) {
        for (i32 i = 0; i < c.nc; ++i) sgs_row(c, s, i, r, z);
        for (i32 i = c.nc - 1; i >= 0; --i) sgs_row(c, s, i, r, z);
    } else {
        for (i32 k = 0; k < s.n_colours; ++k)
            for (i32 i = 0; i < c.nc; ++i)
                if (s.colour[static_cast<size_t>(i)] == k) sgs_row(c, s, i, r, z);
        for (i32 k = s.n_colours - 1; k >= 0; --k)
            for (i32 i = c.nc - 1; i >= 0; --i)
                if (s.colour[static_cast<size_t>(i)] == k) sgs_row(c, s, i, r, z);
    }
}

double dot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

void axpy(std::vector<double>& y, double alpha, const std::vector<double>& x) {
    for (size_t i = 0; i < y.size(); ++i) y[i] += alpha * x[i];
}

// bicgstab.cpp:42-201
oracle_iter_result bicgstab(const oracle_ctx& c, const Sgs* M, const oracle_linparams& lp, std::vector<double>& x,
                            const std::vector<double>& b) {
    const size_t N = x.size();
    oracle_iter_result res{2, 0, 0.0, 0.0};
    std::vector<double> r(N), r0(N), p(N), v(N), s(N), t(N), ph(N), sh(N);
    const auto precond = [&](const std::vector<double>& in, std::vector<double>& out) {
        if (M) sgs_apply(c, *M, in.data(), out.data());
        else out = in;
    };
    spmv(c, x.data(), v.data());
    r = b;
    axpy(r, -1.0, v);
    const double bnorm = std::sqrt(dot(b, b));
    double rnorm = std::sqrt(dot(r, r));
    const double scale = (bnorm > 0.0) ? bnorm : 1.0;
    const double tol = std::max(lp.abs_tol, lp.rel_tol * scale);
    res.initial_residual = rnorm;
    if (!std::isfinite(rnorm)) {
        res.status = 1;
    } else if (rnorm <= tol) {
        res.status = 0;
    }
    int it = 0;
    if (res.status == 2) {
        r0 = r;
        std::fill(v.begin(), v.end(), 0.0);
        double rho = rnorm * rnorm, rho_old = 1.0, alpha = 1.0, omega = 1.0;
        while (true) {
            if (it >= lp.max_iter) {
                res.status = 2;
                break;
            }
            ++it;
            if (it == 1) {
                p = r;
            } else {
                const double beta = (rho / rho_old) * (alpha / omega);
                if (!std::isfinite(beta)) {
                    res.status = 1;
                    break;
                }
                for (auto& e : p) e *= beta;
                axpy(p, -beta * omega, v);
                axpy(p, 1.0, r);
            }
            precond(p, ph);
            spmv(c, ph.data(), v.data());
            const double gamma = dot(r0, v);
            if (!std::isfinite(gamma) || gamma == 0.0) {
                res.status = 1;
                break;
            }
            alpha = rho / gamma;
            s = r;
            axpy(s, -alpha, v);
            precond(s, sh);
            spmv(c, sh.data(), t.data());
            const double tt = dot(t, t), ts = dot(t, s), r0t = dot(r0, t), r0s = dot(r0, s), ss = dot(s, s);
            if (!std::isfinite(ts) || !std::isfinite(tt) || tt == 0.0) {
                res.status = 1;
                break;
            }
            const double omega_new = ts / tt;
            axpy(x, alpha, ph);
            axpy(x, omega_new, sh);
            r = s;
            axpy(r, -omega_new, t);
            const double rn2 = ss - 2.0 * omega_new * ts + omega_new * omega_new * tt;
            rnorm = std::sqrt(std::max(rn2, 0.0));
            const double rho_new = r0s - omega_new * r0t;
            if (!std::isfinite(rnorm) || !std::isfinite(omega_new) || omega_new == 0.0) {
                res.status = 1;
                break;
            }
            if (it % 32 == 0) rnorm = std::sqrt(dot(r, r));
            if (rnorm <= tol) {
                res.status = 0;
                break;
            }
            if (!std::isfinite(rho_new) || rho_new == 0.0) {
                res.status = 1;
                break;
            }
            rho_old = rho;
            rho = rho_new;
            omega = omega_new;
        }
    }
    res.iterations = it;
    res.final_residual = rnorm;
    return res;
}

BcParams resolve_bc(const Eos& eos, const oracle_bc& d) {  // bc_manager.hpp:235-352
    BcParams p;
    p.type = d.type;
    double pp = 0, u = 0, v = 0, w = 0, T = 0;
    switch (d.type) {
        case SupIn:
        case Far:
            resolve_inflow(eos, d, d.p, pp, u, v, w, T);
            p.p = pp;
            p.u = u;
            p.v = v;
            p.w = w;
            p.T = T;
            break;
        case NoSlip:
            p.u = d.velocity[0];
            p.v = d.velocity[1];
            p.w = d.velocity[2];
            p.T = d.t;
            break;
        case NoSlipHF:
            p.u = d.velocity[0];
            p.v = d.velocity[1];
            p.w = d.velocity[2];
            p.tmp_grad = d.tmp_grad;
            break;
        case SubIn: {
            const double p_ref = (d.p > 0.0) ? d.p : 101325.0;
            resolve_inflow(eos, d, p_ref, pp, u, v, w, T);
            p.u = u;
            p.v = v;
            p.w = w;
            p.T = T;
            break;
        }
        case SubOut:
            p.p = d.p;
            break;
        default:
            break;
    }
    return p;
}

}  // namespace

#define ORACLE_TRY(ctx, body)                  \
    try {                                      \
        body;                                  \
        return 0;                              \
    } catch (const std::exception& e) {        \
        if (ctx) (ctx)->err = e.what();        \
        return -1;                             \
    }

extern "C" {

int oracle_create(const oracle_mesh_view* m, const oracle_numerics* num, const oracle_bc* bcs, int32_t n_bcs,
                  oracle_ctx** out) {
    if (!m || !num || !out || (n_bcs > 0 && !bcs)) return -1;
    *out = nullptr;
    auto* c = new oracle_ctx();
    try {
        c->nc = m->n_cells;
        c->nf = m->n_faces;
        c->ninner = m->n_inner_faces;
        c->nbf = c->nf - c->ninner;
        c->ntot = c->nc + c->nbf;
        c->npatch = m->n_patches;
        const size_t nc = static_cast<size_t>(c->nc), nf = static_cast<size_t>(c->nf);
        c->owner.assign(m->face_owner, m->face_owner + nf);
        c->neigh.assign(m->face_neigh, m->face_neigh + nf);
        c->patch.assign(m->face_patch, m->face_patch + nf);
        c->cell_nodes_off.assign(m->cell_nodes_off, m->cell_nodes_off + nc + 1);
        c->cell_nodes.assign(m->cell_nodes, m->cell_nodes + m->cell_nodes_off[nc]);
        c->cx.assign(m->cell_cx, m->cell_cx + nc);
        c->cy.assign(m->cell_cy, m->cell_cy + nc);
        c->cz.assign(m->cell_cz, m->cell_cz + nc);
        c->vol.assign(m->cell_vol, m->cell_vol + nc);
        c->fx.assign(m->face_cx, m->face_cx + nf);
        c->fy.assign(m->face_cy, m->face_cy + nf);
        c->fz.assign(m->face_cz, m->face_cz + nf);
        c->nx.assign(m->face_nx, m->face_nx + nf);
        c->ny.assign(m->face_ny, m->face_ny + nf);
        c->nz.assign(m->face_nz, m->face_nz + nf);
        c->area.assign(m->face_area, m->face_area + nf);
        c->num = *num;
        c->eos = Eos{num->gamma, num->gas_constant};
        if (n_bcs != c->npatch) throw std::runtime_error("one boundary condition per patch required");
        // patch ranges (bc_manager.hpp:336-357), instantiated in patch-id order
        for (i32 pid = 0; pid < n_bcs; ++pid) {
            BcParams bp = resolve_bc(c->eos, bcs[pid]);
            i32 b = -1, e = -1;
            for (i32 f = c->ninner; f < c->nf; ++f) {
                if (c->patch[static_cast<size_t>(f)] == pid) {
                    if (b < 0) b = f;
                    e = f + 1;
                }
            }
            if (b < 0) continue;
            bp.fbeg = b;
            bp.fend = e;
            c->bcs.push_back(bp);
        }
        c->n_nodes = m->n_nodes;
        c->patch_type.assign(static_cast<size_t>(c->npatch), -1);
        for (i32 p = 0; p < n_bcs && p < c->npatch; ++p) c->patch_type[static_cast<size_t>(p)] = bcs[p].type;
        build_cell_faces(*c);
        build_cell_cells_face(*c);
        if (num->gradient == 2) {
            build_cell_cells_node(*c, m->n_nodes);
            setup_lsq(*c);
        }
        setup_face_distances(*c);
        build_bsr_pattern(*c);
        const size_t nt = static_cast<size_t>(c->ntot);
        c->U.assign(5 * nt, 0.0);
        c->q.assign(5 * nt, 0.0);
        c->gx.assign(5 * nt, 0.0);
        c->gy.assign(5 * nt, 0.0);
        c->gz.assign(5 * nt, 0.0);
        c->phi.assign(5 * nt, 0.0);
        c->res.assign(5 * nt, 0.0);
        c->fscale.assign(5 * nt, 0.0);
        c->bflux.assign(5 * static_cast<size_t>(c->nf - c->ninner), 0.0);
        c->lam.assign(nc, 0.0);
        c->dt.assign(nc, 0.0);
        c->alpha.assign(nc, 0.0);
        c->du.assign(5 * nc, 0.0);
    } catch (const std::exception& e) {
        delete c;
        return -1;
    }
    *out = c;
    return 0;
}

int oracle_destroy(oracle_ctx* c) {
    delete c;
    return 0;
}

const char* oracle_last_error(const oracle_ctx* c) { return c ? c->err.c_str() : "null context"; }

int oracle_sizes(const oracle_ctx* c, int32_t* n_cells, int32_t* n_total, int64_t* nnzb) {
    if (!c) return -1;
    if (n_cells) *n_cells = c->nc;
    if (n_total) *n_total = c->ntot;
    if (nnzb) *nnzb = static_cast<int64_t>(c->cols.size());
    return 0;
}

int oracle_set_state(oracle_ctx* c, const double* U, int64_t ld) {
    if (!c || !U || ld < c->nc) return -1;
    for (int v = 0; v < 5; ++v) std::copy_n(U + v * ld, c->nc, plane(c->U, c->ntot, v));
    c->have_residual = false;
    return 0;
}

int oracle_get_state(oracle_ctx* c, double* U, int64_t ld) {
    if (!c || !U || ld < c->nc) return -1;
    for (int v = 0; v < 5; ++v) std::copy_n(plane(c->U, c->ntot, v), c->nc, U + v * ld);
    return 0;
}

// solver.hpp:507-543 with primitives_rhop_to_conserved (state_conversions.hpp:95-111)
int oracle_init_uniform(oracle_ctx* c, double rho, const double vel[3], double p) {
    if (!c || !vel) return -1;
    const double T = c->eos.temperature_rhop(rho, p);
    const double e = c->eos.internal_energy_Tp(T);
    const double ke = 0.5 * rho * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
    const double Uv[5] = {rho, rho * vel[0], rho * vel[1], rho * vel[2], rho * e + ke};
    for (int v = 0; v < 5; ++v) std::fill_n(plane(c->U, c->ntot, v), c->ntot, Uv[v]);
    c->have_residual = false;
    return 0;
}

int oracle_eval_residual(oracle_ctx* c) { ORACLE_TRY(c, evaluate_residual(*c)) }

int oracle_get_primitives(oracle_ctx* c, double* q, int64_t ld) {
    if (!c || !q || ld < c->ntot) return -1;
    for (int v = 0; v < 5; ++v) std::copy_n(plane(c->q, c->ntot, v), c->ntot, q + v * ld);
    return 0;
}

int oracle_get_gradients(oracle_ctx* c, double* g, int64_t ld, double* phi, int64_t ld_phi) {
    if (!c || ld < c->ntot) return -1;
    if (g) {
        for (int v = 0; v < 5; ++v) {
            std::copy_n(plane(c->gx, c->ntot, v), c->ntot, g + (3 * v + 0) * ld);
            std::copy_n(plane(c->gy, c->ntot, v), c->ntot, g + (3 * v + 1) * ld);
            std::copy_n(plane(c->gz, c->ntot, v), c->ntot, g + (3 * v + 2) * ld);
        }
    }
    if (phi) {
        if (ld_phi < c->nc) return -1;
        for (int v = 0; v < 5; ++v) std::copy_n(plane(c->phi, c->ntot, v), c->nc, phi + v * ld_phi);
    }
    return 0;
}

int oracle_get_residual(oracle_ctx* c, double* R, int64_t ld, double* lam) {
    if (!c || ld < c->nc) return -1;
    if (R)
        for (int v = 0; v < 5; ++v) std::copy_n(plane(c->res, c->ntot, v), c->nc, R + v * ld);
    if (lam) std::copy_n(c->lam.data(), c->nc, lam);
    return 0;
}

int oracle_get_flux_scale(oracle_ctx* c, double* S, int64_t ld) {
    if (!c || !S || ld < c->nc) return -1;
    for (int v = 0; v < 5; ++v) std::copy_n(plane(c->fscale, c->ntot, v), c->nc, S + v * ld);
    return 0;
}

// solver.hpp:740-756
int oracle_residual_norms(oracle_ctx* c, double l2[5]) {
    if (!c || !l2) return -1;
    double loc[5] = {0, 0, 0, 0, 0};
    for (int v = 0; v < 5; ++v) {
        const double* r = plane(c->res, c->ntot, v);
        for (i32 i = 0; i < c->nc; ++i) loc[v] += r[i] * r[i];
    }
    const double scale = 1.0 / static_cast<double>(std::max<i32>(c->nc, 1));
    for (int v = 0; v < 5; ++v) l2[v] = std::sqrt(loc[v] * scale);
    return 0;
}

int oracle_compute_dt(oracle_ctx* c) { ORACLE_TRY(c, compute_dt(*c)) }

int oracle_get_dt(oracle_ctx* c, double* dt, double* alpha) {
    if (!c) return -1;
    if (dt) std::copy(c->dt.begin(), c->dt.end(), dt);
    if (alpha) std::copy(c->alpha.begin(), c->alpha.end(), alpha);
    return 0;
}

int oracle_assemble_jacobian(oracle_ctx* c) { ORACLE_TRY(c, assemble_jacobian(*c)) }

int oracle_get_bsr_pattern(oracle_ctx* c, int32_t* row_ptr, int32_t* cols, int32_t* diag) {
    if (!c) return -1;
    if (row_ptr) std::copy(c->row_ptr.begin(), c->row_ptr.end(), row_ptr);
    if (cols) std::copy(c->cols.begin(), c->cols.end(), cols);
    if (diag) std::copy(c->diag.begin(), c->diag.end(), diag);
    return 0;
}

int oracle_get_bsr_values(oracle_ctx* c, double* v) {
    if (!c || !v) return -1;
    std::copy(c->values.begin(), c->values.end(), v);
    return 0;
}

int oracle_set_bsr_values(oracle_ctx* c, const double* v) {
    if (!c || !v) return -1;
    std::copy_n(v, c->values.size(), c->values.begin());
    return 0;
}

int oracle_linear_solve(oracle_ctx* c, const oracle_linparams* lp, const double* b, const double* x0, double* x,
                        oracle_iter_result* res) {
    if (!c || !lp) return -1;
    try {
        const size_t N = static_cast<size_t>(c->nc) * 5;
        std::vector<double> bb(N), xx(N, 0.0);
        if (b) {
            std::copy_n(b, N, bb.begin());
        } else {
            for (i32 i = 0; i < c->nc; ++i)
                for (int v = 0; v < 5; ++v) bb[static_cast<size_t>(i) * 5 + static_cast<size_t>(v)] = -plane(c->res, c->ntot, v)[i];
        }
        if (x0) std::copy_n(x0, N, xx.begin());
        Sgs sgs;
        Sgs* M = nullptr;
        if (lp->preconditioner == 1) {
            if (!sgs_setup(*c, sgs)) throw std::runtime_error("singular diagonal block");
            sgs.order = (lp->sgs_order == 1) ? 1 : 0;
            if (sgs.order == 0) colour_rows(*c, sgs.colour, sgs.n_colours);
            M = &sgs;
        }
        const oracle_iter_result r = bicgstab(*c, M, *lp, xx, bb);
        c->du = xx;
        if (x) std::copy(xx.begin(), xx.end(), x);
        if (res) *res = r;
        return 0;
    } catch (const std::exception& e) {
        c->err = e.what();
        return -1;
    }
}

// implicit_euler.hpp:34-92
int oracle_implicit_step(oracle_ctx* c, const oracle_linparams* lp, oracle_step_result* out) {
    if (!c || !lp) return -1;
    try {
        evaluate_residual(*c);
        compute_dt(*c);
        assemble_jacobian(*c);
        oracle_iter_result lr{};
        if (oracle_linear_solve(c, lp, nullptr, nullptr, nullptr, &lr) != 0) return -1;
        double l2[5];
        oracle_residual_norms(c, l2);
        for (i32 i = 0; i < c->nc; ++i)
            for (int v = 0; v < 5; ++v) plane(c->U, c->ntot, v)[i] = plane(c->U, c->ntot, v)[i] + c->du[static_cast<size_t>(i) * 5 + static_cast<size_t>(v)];
        if (c->sa_on()) sa_update(*c);
        if (out) {
            for (int v = 0; v < 5; ++v) out->l2[v] = l2[v];
            out->linear = lr;
        }
        return 0;
    } catch (const std::exception& e) {
        c->err = e.what();
        return -1;
    }
}

// ForwardEuler::advance (forward_euler.hpp:24-30): residual, explicit dt,
// U <- U - alpha R (update_ops.hpp:28-44); l2 = norms of that residual.
int oracle_explicit_step(oracle_ctx* c, double l2[5]) {
    if (!c) return -1;
    try {
        if (c->sa_on()) throw std::runtime_error("explicit stepping with the turbulence module is not supported");
        if (c->bdf_order > 0) throw std::runtime_error("explicit stepping in dual time is not supported");
        evaluate_residual(*c);
        compute_dt_explicit(*c);
        if (l2) oracle_residual_norms(c, l2);
        for (int v = 0; v < 5; ++v) {
            double* u = plane(c->U, c->ntot, v);
            const double* r = plane(c->res, c->ntot, v);
            for (i32 i = 0; i < c->nc; ++i) u[i] = u[i] - c->alpha[static_cast<size_t>(i)] * r[i];
        }
        return 0;
    } catch (const std::exception& e) {
        c->err = e.what();
        return -1;
    }
}

// Dual time stepping on (order 1 or 2; 0 = steady): U^n = U^{n-1} = the
// current state, as Solver::init_fields sets them (solver.hpp:494-528).
int oracle_set_unsteady(oracle_ctx* c, double dt, int32_t order) {
    if (!c) return -1;
    if (order < 0 || order > 2 || (order > 0 && !(dt > 0.0))) {
        c->err = "unsteady: order 0, 1 or 2 and dt > 0 required";
        return -1;
    }
    if (order > 0 && c->sa_on() && c->sa.cfg.update != 0) {
        c->err = "dual time stepping supports the turbulence module only with the frozen update";
        return -1;
    }
    c->bdf_order = order;
    c->phys_dt = dt;
    c->bdf_cur = 1;
    c->phys_step = 0;
    const size_t n = static_cast<size_t>(c->nc);
    c->Un.assign(5 * n, 0.0);
    for (int v = 0; v < 5; ++v) std::copy_n(plane(c->U, c->ntot, v), n, c->Un.data() + static_cast<size_t>(v) * n);
    c->Unm1 = c->Un;
    return 0;
}

// Solver::run_dual_time (solver.hpp:662-736): per physical step the BDF order
// (1 on the first step since dual time was set; the count carries over
// between calls), implicit sub-iterations with the relative residual checked
// at sub-iteration 1, every 5th and the last, then the time levels shift.
// subiters receives the sub-iterations of every physical step.
int oracle_run_unsteady(oracle_ctx* c, const oracle_linparams* lp, int64_t steps, int64_t max_subiter, double sub_tol,
                        int64_t* subiters) {
    if (!c || !lp) return -1;
    if (c->bdf_order == 0) {
        c->err = "unsteady: not enabled";
        return -1;
    }
    const size_t n = static_cast<size_t>(c->nc);
    for (int64_t k = 0; k < steps; ++k) {
        const int64_t step = ++c->phys_step;  // counts on across calls
        c->bdf_cur = (step == 1 || c->bdf_order == 1) ? 1 : 2;
        double norm0[5] = {0, 0, 0, 0, 0};
        bool have0 = false;
        int64_t last = 0;
        for (int64_t it = 1; it <= max_subiter; ++it) {
            last = it;
            oracle_step_result st{};
            if (oracle_implicit_step(c, lp, &st) != 0) return -1;
            const bool diagnose = it == 1 || it % 5 == 0 || it == max_subiter;
            if (!diagnose) continue;
            if (!std::isfinite(st.l2[0])) {
                c->err = "unsteady: non-finite residual";
                return -1;
            }
            double rel = 1.0;
            if (!have0) {
                for (int v = 0; v < 5; ++v) norm0[v] = st.l2[v];
                have0 = true;
            } else {
                rel = 0.0;
                for (int v = 0; v < 5; ++v) rel = std::max(rel, st.l2[v] / std::max(norm0[v], 1.0e-300));
            }
            if (rel <= sub_tol) break;
        }
        if (subiters) subiters[k] = last;
        // U^{n-1} = U^n, U^n = U^{n+1}
        if (c->bdf_order == 2) std::swap(c->Unm1, c->Un);
        for (int v = 0; v < 5; ++v) std::copy_n(plane(c->U, c->ntot, v), n, c->Un.data() + static_cast<size_t>(v) * n);
    }
    return 0;
}

int oracle_row_groups(oracle_ctx* c, int32_t* colour, int32_t* n_colours, int32_t* fwd, int32_t* n_fwd, int32_t* bwd,
                      int32_t* n_bwd) {
    if (!c) return -1;
    std::vector<i32> col;
    i32 nco = 0;
    colour_rows(*c, col, nco);
    if (colour) std::copy(col.begin(), col.end(), colour);
    if (n_colours) *n_colours = nco;
    std::vector<i32> lf(static_cast<size_t>(c->nc), 0), lb(static_cast<size_t>(c->nc), 0);
    i32 mf = 0, mb = 0;
    for (i32 i = 0; i < c->nc; ++i) {
        i32 l = 0;
        for (i32 e = c->ccf_off[static_cast<size_t>(i)]; e < c->ccf_off[static_cast<size_t>(i) + 1]; ++e) {
            const i32 j = c->ccf[static_cast<size_t>(e)];
            if (j < i) l = std::max(l, lf[static_cast<size_t>(j)] + 1);
        }
        lf[static_cast<size_t>(i)] = l;
        mf = std::max(mf, l + 1);
    }
    for (i32 i = c->nc - 1; i >= 0; --i) {
        i32 l = 0;
        for (i32 e = c->ccf_off[static_cast<size_t>(i)]; e < c->ccf_off[static_cast<size_t>(i) + 1]; ++e) {
            const i32 j = c->ccf[static_cast<size_t>(e)];
            if (j > i) l = std::max(l, lb[static_cast<size_t>(j)] + 1);
        }
        lb[static_cast<size_t>(i)] = l;
        mb = std::max(mb, l + 1);
    }
    if (fwd) std::copy(lf.begin(), lf.end(), fwd);
    if (bwd) std::copy(lb.begin(), lb.end(), bwd);
    if (n_fwd) *n_fwd = mf;
    if (n_bwd) *n_bwd = mb;
    return 0;
}

int oracle_sgs_apply(oracle_ctx* c, int32_t sgs_order, const double* r, double* z) {
    if (!c || !r || !z) return -1;
    Sgs s;
    if (!sgs_setup(*c, s)) {
        c->err = "singular diagonal block";
        return -1;
    }
    s.order = (sgs_order == 1) ? 1 : 0;
    if (s.order == 0) colour_rows(*c, s.colour, s.n_colours);
    sgs_apply(*c, s, r, z);
    return 0;
}

// ResidualKernel::boundary_integrals (residual_kernel.hpp:96-169): per patch,
// in ascending face order, the net mass flux F[0] and the energy flux F[4]
// (inviscid + viscous) of the boundary faces.  force[3p + k] adds F[1 + k],
// the momentum flux out of the domain (the force of the fluid on the patch),
// which the reference does not form.
int oracle_boundary_integrals(oracle_ctx* c, double* mass, double* energy, double* force, int32_t n_patches) {
    if (!c) return -1;
    if (!c->have_residual) {
        c->err = "boundary_integrals needs a residual";
        return -3;
    }
    if (n_patches != c->npatch) {
        c->err = "n_patches mismatch";
        return -1;
    }
    for (i32 p = 0; p < n_patches; ++p) {
        if (mass) mass[p] = 0.0;
        if (energy) energy[p] = 0.0;
        if (force) force[3 * p] = force[3 * p + 1] = force[3 * p + 2] = 0.0;
    }
    for (i32 f = c->ninner; f < c->nf; ++f) {
        const size_t b = static_cast<size_t>(f - c->ninner) * 5;
        const i32 p = c->patch[static_cast<size_t>(f)];
        if (mass) mass[p] += c->bflux[b];
        if (energy) energy[p] += c->bflux[b + 4];
        if (force)
            for (int k = 0; k < 3; ++k) force[3 * p + k] += c->bflux[b + 1 + static_cast<size_t>(k)];
    }
    return 0;
}

// MeshAuxGeometry::add_wall_distance + relax_distance_graph
// (aux_geometry.cpp:545-697): cells owning a no-slip wall face start at the
// distance to its centroid, every other cell at 1e30; then forward and backward
// Gauss-Seidel sweeps d_c = min(d_c, d_n + |x_n - x_c|) over the face (stencil
// 0) or vertex (stencil 1) neighbours until the largest relative decrease is
// <= rel_tol or max_sweeps sweeps have run.
int oracle_wall_distance(oracle_ctx* c, int32_t stencil, int32_t max_sweeps, double rel_tol, double* dist) {
    if (!c || !dist) return -1;
    wall_distance(*c, stencil, max_sweeps, rel_tol, dist);
    return 0;
}

// Prescribed eddy viscosity of every cell and ghost (n_total values), read by
// the mean viscous flux (residual_kernel.hpp:153-154, 231-232, 284-285) and the
// Jacobian (jacobian_kernel.hpp:156-157, 209-210, 252-254: ghost side zero);
// mut == NULL clears it.
int oracle_set_eddy_viscosity(oracle_ctx* c, const double* mut, int32_t n_total) {
    if (!c) return -1;
    if (!mut) {
        c->mut_ext.clear();
        c->have_residual = false;
        return 0;
    }
    if (c->sa_on()) {
        c->err = "the turbulence module forms its own eddy viscosity";
        return -3;
    }
    if (n_total != c->ntot || c->num.flow_model != 1) {
        c->err = "eddy viscosity needs n_total values and the viscous flow model";
        return -1;
    }
    c->mut_ext.assign(mut, mut + n_total);
    c->have_residual = false;
    return 0;
}

int oracle_set_turbulence(oracle_ctx* c, const oracle_turbulence* t) {
    if (!c || !t) return -1;
    if (t->model != 0) c->mut_ext.clear();
    if (t->model != 0 && c->bdf_order > 0 && t->update != 0) {
        c->err = "dual time stepping supports the turbulence module only with the frozen update";
        return -1;
    }
    if (t->model == 0) {
        c->sa.cfg = *t;
        return 0;
    }
    if (t->model != 1 || (t->update != 0 && t->update != 1) || c->num.flow_model != 1 || !(t->nu_inf_ratio > 0.0) ||
        !(t->rho_inf > 0.0) || !(t->p_inf > 0.0) || t->max_distance_sweeps < 0) {
        c->err = "invalid turbulence parameters (SA needs the viscous flow model)";
        return -1;
    }
    auto& s = c->sa;
    s.cfg = *t;
    const size_t nt = static_cast<size_t>(c->ntot), nc = static_cast<size_t>(c->nc);
    s.kind.assign(static_cast<size_t>(c->nbf), ZSymmetry);
    for (i32 f = c->ninner; f < c->nf; ++f) {
        const i32 p = c->patch[static_cast<size_t>(f)];
        const int ty = (p >= 0 && p < c->npatch) ? c->patch_type[static_cast<size_t>(p)] : Sym;
        s.kind[static_cast<size_t>(f - c->ninner)] = nu_zone(ty);
    }
    // initialize (spalart_allmaras.hpp:157-193): wall distance over the vertex
    // stencil, freestream working values
    s.dist.assign(nc, 0.0);
    wall_distance(*c, 1, t->max_distance_sweeps, t->distance_tolerance, s.dist.data());
    const double T_inf = c->eos.temperature_rhop(t->rho_inf, t->p_inf);
    s.nu_inf = t->nu_inf_ratio * viscosity(T_inf) / t->rho_inf;
    s.rho_nu_inf = t->rho_inf * s.nu_inf;
    s.rnu.assign(nt, s.rho_nu_inf);  // init_state (spalart_allmaras.hpp:184-192)
    s.nu.assign(nt, 0.0);
    s.rho.assign(nt, 0.0);
    s.mut.assign(nt, 0.0);
    s.g.assign(3 * nt, 0.0);
    s.res.assign(nt, 0.0);
    s.diag.assign(nt, 0.0);
    s.mdot.assign(static_cast<size_t>(c->nf), 0.0);
    return 0;
}

int oracle_set_turbulence_state(oracle_ctx* c, const double* rnu) {
    if (!c || !rnu || !c->sa_on()) return -1;
    std::copy(rnu, rnu + c->nc, c->sa.rnu.begin());
    return 0;
}

int oracle_get_turbulence(oracle_ctx* c, double* rnu, double* nu, double* mut, double* dist, double* res, double* grad,
                          double* mdot, double* diag) {
    if (!c || !c->sa_on()) return -1;
    const size_t nc = static_cast<size_t>(c->nc), nt = static_cast<size_t>(c->ntot);
    const auto out = [&](const std::vector<double>& a, double* o) {
        if (o) std::copy(a.begin(), a.begin() + static_cast<long>(nc), o);
    };
    out(c->sa.rnu, rnu);
    out(c->sa.nu, nu);
    out(c->sa.mut, mut);
    out(c->sa.dist, dist);
    out(c->sa.res, res);
    out(c->sa.diag, diag);
    if (grad)
        for (size_t d = 0; d < 3; ++d) std::copy(c->sa.g.begin() + static_cast<long>(d * nt), c->sa.g.begin() + static_cast<long>(d * nt + nc), grad + d * nc);
    if (mdot) std::copy(c->sa.mdot.begin(), c->sa.mdot.end(), mdot);
    return 0;
}

int oracle_sa_point(double rho, double nu, double mu_lam, double d, double S, double grad2, double out[8]) {
    if (!out) return -1;
    const SaSource o = sa_source(rho, nu, mu_lam, d, S, grad2);
    const double v[8] = {o.P, o.D, o.cb2, o.fv1, o.fv2, o.sbar, o.r, o.fw};
    std::copy(v, v + 8, out);
    return 0;
}

double oracle_sa_constant(int32_t which) {
    const double v[8] = {kSaCb1, kSaCb2, kSaSigma, kSaKappa, kSaCw1, kSaCw2, kSaCw3, kSaCv1};
    return (which >= 0 && which < 8) ? v[which] : 0.0;
}

}  // extern "C"

namespace {

// aux_geometry.cpp:545-697 (single rank): seeds at the no-slip wall faces, then
// forward + backward relaxation sweeps over the face (0) or vertex (1) stencil
void wall_distance(oracle_ctx& cr, int32_t stencil, int32_t max_sweeps, double rel_tol, double* dist) {
    oracle_ctx* c = &cr;
    const double kInf = 1.0e30, kMinDist = 1.0e-12;
    std::vector<double> d(static_cast<size_t>(c->nc), kInf);
    for (i32 f = c->ninner; f < c->nf; ++f) {
        const size_t fs = static_cast<size_t>(f);
        const int t = c->patch_type[static_cast<size_t>(c->patch[fs])];
        if (t != 5 && t != 6) continue;  // NO_SLIP_WALL, NO_SLIP_WALL_HEAT_FLUX
        const size_t c0 = static_cast<size_t>(c->owner[fs]);
        const double ex = c->fx[fs] - c->cx[c0], ey = c->fy[fs] - c->cy[c0], ez = c->fz[fs] - c->cz[c0];
        d[c0] = std::min(d[c0], std::sqrt(ex * ex + ey * ey + ez * ez));
    }
    if (stencil == 1 && c->ccn_off.empty()) build_cell_cells_node(*c, c->n_nodes);
    const std::vector<i32>& off = (stencil == 1) ? c->ccn_off : c->ccf_off;
    const std::vector<i32>& nb = (stencil == 1) ? c->ccn : c->ccf;
    std::vector<double> ed(nb.size());
    for (i32 i = 0; i < c->nc; ++i) {
        const size_t is = static_cast<size_t>(i);
        for (i32 j = off[is]; j < off[is + 1]; ++j) {
            const size_t n = static_cast<size_t>(nb[static_cast<size_t>(j)]);
            const double dx = c->cx[n] - c->cx[is], dy = c->cy[n] - c->cy[is], dz = c->cz[n] - c->cz[is];
            ed[static_cast<size_t>(j)] = std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    }
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
        double max_rel = 0.0;
        const auto relax = [&](size_t cell) {
            const double d0 = d[cell];
            double dn = d0;
            for (i32 j = off[cell]; j < off[cell + 1]; ++j) {
                const double cand = d[static_cast<size_t>(nb[static_cast<size_t>(j)])] + ed[static_cast<size_t>(j)];
                dn = (cand < dn) ? cand : dn;
            }
            if (dn < d0) {
                const double rel = (d0 - dn) / std::max(d0, kMinDist);
                max_rel = (rel > max_rel) ? rel : max_rel;
                d[cell] = dn;
            }
        };
        for (size_t cell = 0; cell < static_cast<size_t>(c->nc); ++cell) relax(cell);
        for (size_t cell = static_cast<size_t>(c->nc); cell-- > 0;) relax(cell);
        if (max_rel <= rel_tol) break;
    }
    std::copy(d.begin(), d.end(), dist);
}

}  // namespace

extern "C" {

int oracle_spmv(oracle_ctx* c, const double* x, double* y) {
    if (!c || !x || !y) return -1;
    spmv(*c, x, y);
    return 0;
}

}  // extern "C"
