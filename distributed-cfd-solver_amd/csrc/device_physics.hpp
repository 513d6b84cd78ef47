// Per-face / per-cell physics evaluated on the device (and usable on the host).
//
// Every formula keeps the reference's operation order so that, with FP
// contraction disabled, a device result equals the CPU result bit for bit
// wherever the operations are IEEE-exact (+, -, *, /, sqrt):
//   ideal gas and state conversions   eos/ideal_gas.hpp:13-85, state_conversions.hpp:12-129
//   HLLC flux, Picard Jacobian        fluxes/hllc.hpp:22-274
//   viscous flux and its Jacobian     fluxes/viscous.hpp:68-359
//   slope limiters                    numerics/limiter/slope_limiters.hpp:21-92
//   farfield Riemann state            bc/physical/farfield.hpp:39-113
#pragma once

#include <hip/hip_runtime.h>

#define CFDG_HD __host__ __device__ __forceinline__

namespace cfdg {

struct Gas {
    double gamma;
    double R;
    double prandtl;
    double prandtl_turb;
    double cv, cp;  // R / (gamma - 1) and gamma cv, formed once by make_gas
};

// The gas with its specific heats formed once (the same IEEE operations on
// the host as on the device, so kernels read them from their arguments
// instead of forming them per thread).
CFDG_HD Gas make_gas(double gamma, double R, double prandtl, double prandtl_turb) {
    const double cv = R / (gamma - 1.0);
    return Gas{gamma, R, prandtl, prandtl_turb, cv, gamma * cv};
}
CFDG_HD double gas_cv(const Gas& g) { return g.cv; }
CFDG_HD double gas_cp(const Gas& g) { return g.cp; }

CFDG_HD double dmax(double a, double b) { return (a < b) ? b : a; }  // std::max semantics
CFDG_HD double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min semantics

// e = E/rho - |u|^2/2 with |u|^2 from the momenta
CFDG_HD double internal_energy_cons(const double U[5]) {
    const double inv_rho = 1.0 / U[0];
    const double q2 = (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) * (inv_rho * inv_rho);
    return (U[4] * inv_rho) - 0.5 * q2;
}

CFDG_HD double pressure_cons(const Gas& g, const double U[5]) {
    return (g.gamma - 1.0) * U[0] * internal_energy_cons(U);
}

// conservative -> (p, u, v, w, T)
CFDG_HD void cons_to_prim(const Gas& g, const double U[5], double q[5]) {
    const double rho = U[0];
    const double inv_rho = 1.0 / rho;
    const double u = U[1] * inv_rho;
    const double v = U[2] * inv_rho;
    const double w = U[3] * inv_rho;
    const double q2 = u * u + v * v + w * w;
    const double e = (U[4] * inv_rho) - 0.5 * q2;
    q[0] = (g.gamma - 1.0) * rho * e;
    q[1] = u;
    q[2] = v;
    q[3] = w;
    q[4] = e / gas_cv(g);
}

// (p, u, v, w, T) -> conservative
CFDG_HD void prim_to_cons(const Gas& g, const double q[5], double U[5]) {
    const double rho = q[0] / (g.R * q[4]);
    const double e = gas_cv(g) * q[4];
    const double ke = 0.5 * rho * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    U[0] = rho;
    U[1] = rho * q[1];
    U[2] = rho * q[2];
    U[3] = rho * q[3];
    U[4] = rho * e + ke;
}

// HLLC star-region flux of side K (contact-resolving branch)
CFDG_HD void hllc_star(const double UK[5], const double FK[5], double rhoK, double pK, double ux, double uy,
                       double uz, double unk, double ak, double SK, double SM, double nx, double ny, double nz,
                       double area, double F[5]) {
    const double den = SK - SM;
    if (fabs(den) < 1.0e-6 * (ak + fabs(SK))) {
        for (int v = 0; v < 5; ++v) F[v] = FK[v] * area;
        return;
    }
    const double inv_den = 1.0 / den;
    const double d_wave = SK - unk;
    const double qf = d_wave * inv_den;
    const double d = SM - unk;
    const double p_term = pK / (rhoK * d_wave);
    const double e_star = UK[4] + rhoK * d * (SM + p_term);
    const double Ust[5] = {qf * rhoK, qf * rhoK * (ux + d * nx), qf * rhoK * (uy + d * ny), qf * rhoK * (uz + d * nz),
                           qf * e_star};
    for (int v = 0; v < 5; ++v) F[v] = (FK[v] + SK * (Ust[v] - UK[v])) * area;
}

// HLLC with Roe-averaged Einfeldt/Davis bounds; F already multiplied by the area.
CFDG_HD void hllc_flux(const Gas& g, const double UL[5], const double UR[5], double nx, double ny, double nz,
                       double area, double F[5], double& smax) {
    const double rhoL = UL[0], rhoR = UR[0];
    const double inv_rhoL = 1.0 / rhoL, inv_rhoR = 1.0 / rhoR;
    const double pL = pressure_cons(g, UL), pR = pressure_cons(g, UR);
    const double uLx = UL[1] * inv_rhoL, uLy = UL[2] * inv_rhoL, uLz = UL[3] * inv_rhoL;
    const double uRx = UR[1] * inv_rhoR, uRy = UR[2] * inv_rhoR, uRz = UR[3] * inv_rhoR;
    const double unL = uLx * nx + uLy * ny + uLz * nz;
    const double unR = uRx * nx + uRy * ny + uRz * nz;
    const double aL = sqrt(g.gamma * pL / rhoL);
    const double aR = sqrt(g.gamma * pR / rhoR);
    const double rl = sqrt(rhoL), rr = sqrt(rhoR);
    const double wr = 1.0 / (rl + rr);
    const double rux = (rl * uLx + rr * uRx) * wr;
    const double ruy = (rl * uLy + rr * uRy) * wr;
    const double ruz = (rl * uLz + rr * uRz) * wr;
    const double HL = (UL[4] + pL) * inv_rhoL;
    const double HR = (UR[4] + pR) * inv_rhoR;
    const double rh = (rl * HL + rr * HR) * wr;
    const double q2 = rux * rux + ruy * ruy + ruz * ruz;
    const double a2 = (g.gamma - 1.0) * (rh - 0.5 * q2);
    const double ra = sqrt(a2 > 0.0 ? a2 : 0.0);
    const double run = rux * nx + ruy * ny + ruz * nz;
    const double SL = dmin(unL - aL, run - ra);
    const double SR = dmax(unR + aR, run + ra);
    smax = dmax(fabs(SL), fabs(SR));
    const double dL = SL - unL, dR = SR - unR;
    const double denom = rhoL * dL - rhoR * dR;
    const double SM = (fabs(denom) > 1e-14) ? (pR - pL + rhoL * unL * dL - rhoR * unR * dR) / denom : 0.5 * (unL + unR);
    if (SM >= 0.0) {
        const double FL[5] = {rhoL * unL, rhoL * uLx * unL + pL * nx, rhoL * uLy * unL + pL * ny,
                              rhoL * uLz * unL + pL * nz, (UL[4] + pL) * unL};
        if (SL >= 0.0) {
            for (int v = 0; v < 5; ++v) F[v] = FL[v] * area;
        } else {
            hllc_star(UL, FL, rhoL, pL, uLx, uLy, uLz, unL, aL, SL, SM, nx, ny, nz, area, F);
        }
    } else {
        const double FR[5] = {rhoR * unR, rhoR * uRx * unR + pR * nx, rhoR * uRy * unR + pR * ny,
                              rhoR * uRz * unR + pR * nz, (UR[4] + pR) * unR};
        if (SR <= 0.0) {
            for (int v = 0; v < 5; ++v) F[v] = FR[v] * area;
        } else {
            hllc_star(UR, FR, rhoR, pR, uRx, uRy, uRz, unR, aR, SR, SM, nx, ny, nz, area, F);
        }
    }
}

// Only the wave-speed bound of hllc_flux (what the Picard Jacobian needs):
// identical arithmetic up to smax, the flux branches are skipped.
CFDG_HD double hllc_smax(const Gas& g, const double UL[5], const double UR[5], double nx, double ny, double nz) {
    const double rhoL = UL[0], rhoR = UR[0];
    const double inv_rhoL = 1.0 / rhoL, inv_rhoR = 1.0 / rhoR;
    const double pL = pressure_cons(g, UL), pR = pressure_cons(g, UR);
    const double uLx = UL[1] * inv_rhoL, uLy = UL[2] * inv_rhoL, uLz = UL[3] * inv_rhoL;
    const double uRx = UR[1] * inv_rhoR, uRy = UR[2] * inv_rhoR, uRz = UR[3] * inv_rhoR;
    const double unL = uLx * nx + uLy * ny + uLz * nz;
    const double unR = uRx * nx + uRy * ny + uRz * nz;
    const double aL = sqrt(g.gamma * pL / rhoL);
    const double aR = sqrt(g.gamma * pR / rhoR);
    const double rl = sqrt(rhoL), rr = sqrt(rhoR);
    const double wr = 1.0 / (rl + rr);
    const double rux = (rl * uLx + rr * uRx) * wr;
    const double ruy = (rl * uLy + rr * uRy) * wr;
    const double ruz = (rl * uLz + rr * uRz) * wr;
    const double HL = (UL[4] + pL) * inv_rhoL;
    const double HR = (UR[4] + pR) * inv_rhoR;
    const double rh = (rl * HL + rr * HR) * wr;
    const double q2 = rux * rux + ruy * ruy + ruz * ruz;
    const double a2 = (g.gamma - 1.0) * (rh - 0.5 * q2);
    const double ra = sqrt(a2 > 0.0 ? a2 : 0.0);
    const double run = rux * nx + ruy * ny + ruz * nz;
    const double SL = dmin(unL - aL, run - ra);
    const double SR = dmax(unR + aR, run + ra);
    return dmax(fabs(SL), fabs(SR));
}

// Analytical Jacobian of the normal Euler flux, row-major 5x5, scaled by s
// after evaluation (J * s), i.e. the reference's "dF *= half_area".
CFDG_HD void euler_jacobian_scaled(const Gas& g, const double U[5], double nx, double ny, double nz, double s,
                                   double J[25]) {
    const double rho = U[0];
    const double inv_rho = 1.0 / rho;
    const double u = U[1] * inv_rho, v = U[2] * inv_rho, w = U[3] * inv_rho, E = U[4] * inv_rho;
    const double p = pressure_cons(g, U);
    const double gamma = g.gamma;
    const double ph = gamma - 1.0;
    const double un = u * nx + v * ny + w * nz;
    const double q2 = u * u + v * v + w * w;
    const double H = E + p * inv_rho;
    const double hpq = 0.5 * ph * q2;
    J[0] = 0.0 * s;
    J[1] = nx * s;
    J[2] = ny * s;
    J[3] = nz * s;
    J[4] = 0.0 * s;
    J[5] = (hpq * nx - u * un) * s;
    J[6] = (un + nx * (1.0 - ph) * u) * s;
    J[7] = (u * ny - ph * nx * v) * s;
    J[8] = (u * nz - ph * nx * w) * s;
    J[9] = (ph * nx) * s;
    J[10] = (hpq * ny - v * un) * s;
    J[11] = (v * nx - ph * ny * u) * s;
    J[12] = (un + ny * (1.0 - ph) * v) * s;
    J[13] = (v * nz - ph * ny * w) * s;
    J[14] = (ph * ny) * s;
    J[15] = (hpq * nz - w * un) * s;
    J[16] = (w * nx - ph * nz * u) * s;
    J[17] = (w * ny - ph * nz * v) * s;
    J[18] = (un + nz * (1.0 - ph) * w) * s;
    J[19] = (ph * nz) * s;
    J[20] = (un * (hpq - H)) * s;
    J[21] = (H * nx - ph * u * un) * s;
    J[22] = (H * ny - ph * v * un) * s;
    J[23] = (H * nz - ph * w * un) * s;
    J[24] = (gamma * un) * s;
}

// Sutherland viscosity
CFDG_HD double sutherland(double T) {
    constexpr double kConst = 1.716e-5 * (273.15 + 110.4);
    const double ratio = T / 273.15;
    return kConst * ratio * sqrt(ratio) / (T + 110.4);
}

// Viscous face flux (already multiplied by the area) and its spectral-radius
// contribution.  qa/qb: (p,u,v,w,T) of the owner and neighbour (or ghost);
// ga/gb: gradients of u, v, w, T, 3 components each ([du, dv, dw, dT]).
// (dvx, dvy, dvz) is the owner->neighbour centroid vector, inv_d its inverse length.
// kEddy = false is the form without eddy viscosity (mutL = mutR = 0): the
// terms 0.5 (0 + 0) and (0 cp) / Pr_t it leaves out are +0, and adding +0 to
// the positive laminar values changes no bit.
template <bool kEddy = true>
CFDG_HD void viscous_flux(const Gas& g, double inv_d, double dvx, double dvy, double dvz, const double qa[5],
                          const double qb[5], const double ga[12], const double gb[12], double mutL, double mutR,
                          double nx, double ny, double nz, double area, double Fv[5], double& lam_visc) {
    const double p_f = 0.5 * (qa[0] + qb[0]);
    const double u_f = 0.5 * (qa[1] + qb[1]);
    const double v_f = 0.5 * (qa[2] + qb[2]);
    const double w_f = 0.5 * (qa[3] + qb[3]);
    const double T_f = 0.5 * (qa[4] + qb[4]);
    const double mu_lam = sutherland(T_f);
    const double k_lam = (sutherland(T_f) * gas_cp(g)) / g.prandtl;
    double mu_eff = mu_lam, k_eff = k_lam;
    if (kEddy) {
        const double mu_t_f = 0.5 * (mutL + mutR);
        mu_eff = mu_lam + mu_t_f;
        const double cp = gas_cp(g);
        k_eff = k_lam + (mu_t_f * cp) / g.prandtl_turb;
    }
    const double xi_x = dvx * inv_d, xi_y = dvy * inv_d, xi_z = dvz * inv_d;
    const double ncx = nx - xi_x, ncy = ny - xi_y, ncz = nz - xi_z;
    const double du_dx = 0.5 * (ga[0] + gb[0]), du_dy = 0.5 * (ga[1] + gb[1]), du_dz = 0.5 * (ga[2] + gb[2]);
    const double dv_dx = 0.5 * (ga[3] + gb[3]), dv_dy = 0.5 * (ga[4] + gb[4]), dv_dz = 0.5 * (ga[5] + gb[5]);
    const double dw_dx = 0.5 * (ga[6] + gb[6]), dw_dy = 0.5 * (ga[7] + gb[7]), dw_dz = 0.5 * (ga[8] + gb[8]);
    const double dT_dx = 0.5 * (ga[9] + gb[9]), dT_dy = 0.5 * (ga[10] + gb[10]), dT_dz = 0.5 * (ga[11] + gb[11]);
    const double div_v = du_dx + dv_dy + dw_dz;
    const double ttd = (2.0 / 3.0) * div_v;
    const double du_dn = (qb[1] - qa[1]) * inv_d + (du_dx * ncx + du_dy * ncy + du_dz * ncz);
    const double dv_dn = (qb[2] - qa[2]) * inv_d + (dv_dx * ncx + dv_dy * ncy + dv_dz * ncz);
    const double dw_dn = (qb[3] - qa[3]) * inv_d + (dw_dx * ncx + dw_dy * ncy + dw_dz * ncz);
    const double dT_dn = (qb[4] - qa[4]) * inv_d + (dT_dx * ncx + dT_dy * ncy + dT_dz * ncz);
    const double tnx = mu_eff * (du_dn + (du_dx * nx + dv_dx * ny + dw_dx * nz) - ttd * nx);
    const double tny = mu_eff * (dv_dn + (du_dy * nx + dv_dy * ny + dw_dy * nz) - ttd * ny);
    const double tnz = mu_eff * (dw_dn + (du_dz * nx + dv_dz * ny + dw_dz * nz) - ttd * nz);
    const double qn = -k_eff * dT_dn;
    const double rho_f = p_f / (g.R * T_f);
    const double nu_eff = mu_eff / rho_f;
    lam_visc = (4.0 / 3.0) * nu_eff * inv_d * area;
    Fv[0] = 0.0;
    Fv[1] = -tnx * area;
    Fv[2] = -tny * area;
    Fv[3] = -tnz * area;
    Fv[4] = (-(u_f * tnx + v_f * tny + w_f * tnz) + qn) * area;
}

// Thin-layer viscous Jacobian dFv/dUL (dFv/dUR = -dFv/dUL) for face-averaged
// primitive state; only the non-zero entries are produced:
//   m[0..2]: row 1..3 first column (-C_mu u/rho, ...), m[3]: C_mu/rho (diagonal of rows 1..3),
//   m[4..8]: row 4 (energy).
// kEddy = false: no eddy viscosity, as in viscous_flux.
template <bool kEddy = true>
CFDG_HD void viscous_jacobian_entries(const Gas& g, double inv_d, const double qL[5], const double qR[5],
                                      double mutL, double mutR, double area, double m[9]) {
    const double p_f = 0.5 * (qL[0] + qR[0]);
    const double u_f = 0.5 * (qL[1] + qR[1]);
    const double v_f = 0.5 * (qL[2] + qR[2]);
    const double w_f = 0.5 * (qL[3] + qR[3]);
    const double T_f = 0.5 * (qL[4] + qR[4]);
    const double rho_f = p_f / (g.R * T_f);
    const double inv_rho = 1.0 / rho_f;
    const double cv = gas_cv(g);
    const double inv_rho_cv = 1.0 / (rho_f * cv);
    const double mu_lam = sutherland(T_f);
    const double k_lam = (sutherland(T_f) * gas_cp(g)) / g.prandtl;
    double mu_eff = mu_lam, k_eff = k_lam;
    if (kEddy) {
        const double mu_t_f = 0.5 * (mutL + mutR);
        mu_eff = mu_lam + mu_t_f;
        const double cp = gas_cp(g);
        k_eff = k_lam + (mu_t_f * cp) / g.prandtl_turb;
    }
    const double C_mu = mu_eff * inv_d * area;
    const double C_lam = k_eff * inv_d * area;
    const double V2 = u_f * u_f + v_f * v_f + w_f * w_f;
    m[0] = -C_mu * u_f * inv_rho;
    m[1] = -C_mu * v_f * inv_rho;
    m[2] = -C_mu * w_f * inv_rho;
    m[3] = C_mu * inv_rho;
    m[4] = C_lam * (0.5 * V2 - cv * T_f) * inv_rho_cv - C_mu * V2 * inv_rho;
    m[5] = -C_lam * u_f * inv_rho_cv + C_mu * u_f * inv_rho;
    m[6] = -C_lam * v_f * inv_rho_cv + C_mu * v_f * inv_rho;
    m[7] = -C_lam * w_f * inv_rho_cv + C_mu * w_f * inv_rho;
    m[8] = C_lam * inv_rho_cv;
}

// Expands viscous_jacobian_entries into the dense row-major pattern.
CFDG_HD void viscous_jacobian_dense(const double m[9], double J[25]) {
    for (int i = 0; i < 25; ++i) J[i] = 0.0;
    J[5] = m[0];
    J[6] = m[3];
    J[10] = m[1];
    J[12] = m[3];
    J[15] = m[2];
    J[18] = m[3];
    J[20] = m[4];
    J[21] = m[5];
    J[22] = m[6];
    J[23] = m[7];
    J[24] = m[8];
}

// Gradient limiter policies phi(delta_nb, delta_face, eps2)
CFDG_HD double limiter_phi(int kind, double d_nb, double d_face, double eps2) {
    if (kind == 3) return 1.0;
    if (d_face == 0.0) return 1.0;
    if (kind == 1) {  // Barth-Jespersen: clamp(r, 0, 1)
        const double r = d_nb / d_face;
        return (r < 0.0) ? 0.0 : ((1.0 < r) ? 1.0 : r);
    }
    const double prod = d_nb * d_face;
    if (prod <= 0.0) return 0.0;
    if (kind == 0) {  // Venkatakrishnan
        const double a2 = d_nb * d_nb + eps2;
        const double num = a2 + 2.0 * prod;
        const double den = a2 + 2.0 * d_face * d_face + prod;
        return num / den;
    }
    // Van Albada
    const double num = d_nb * d_nb + prod;
    const double den = d_nb * d_nb + d_face * d_face + eps2;
    return (den > 0.0) ? (num / den) : 1.0;
}

// Boundary-condition parameters resolved on the host (one entry per patch).
struct BcDev {
    int type;
    double p, u, v, w, T, tmp_grad;
};

enum BcType : int {
    kSupIn = 0, kSupOut = 1, kSubIn = 2, kSubOut = 3, kSlip = 4, kNoSlip = 5, kNoSlipHF = 6, kSym = 7, kFar = 8
};

// Characteristic farfield boundary state (Riemann invariants).
CFDG_HD void farfield_state(const BcDev& p, double gamma, double R, double p_in, double T_in, double vx, double vy,
                            double vz, double nx, double ny, double nz, double qb[5], bool& sup_out) {
    const double gm1 = gamma - 1.0;
    const double inv_gm1 = 1.0 / gm1;
    const double un_in = vx * nx + vy * ny + vz * nz;
    const double un_inf = p.u * nx + p.v * ny + p.w * nz;
    const double a_in = sqrt(gamma * R * dmax(T_in, 1.0e-6));
    const double a_inf = sqrt(gamma * R * dmax(p.T, 1.0e-6));
    const double mn_in = un_in / a_in;
    if (mn_in >= 1.0) {
        sup_out = true;
        qb[0] = p_in;
        qb[1] = vx;
        qb[2] = vy;
        qb[3] = vz;
        qb[4] = T_in;
        return;
    }
    sup_out = false;
    if (mn_in <= -1.0) {
        qb[0] = p.p;
        qb[1] = p.u;
        qb[2] = p.v;
        qb[3] = p.w;
        qb[4] = p.T;
        return;
    }
    const double R_out = un_in + 2.0 * a_in * inv_gm1;
    const double R_in = un_inf - 2.0 * a_inf * inv_gm1;
    const double un_b = 0.5 * (R_out + R_in);
    double a_b = 0.25 * gm1 * (R_out - R_in);
    if (a_b <= 0.0) a_b = a_inf;
    double s;
    if (un_b >= 0.0) {
        qb[1] = vx + (un_b - un_in) * nx;
        qb[2] = vy + (un_b - un_in) * ny;
        qb[3] = vz + (un_b - un_in) * nz;
        const double rho_in = p_in / (R * dmax(T_in, 1.0e-6));
        s = p_in / pow(dmax(rho_in, 1.0e-12), gamma);
    } else {
        qb[1] = p.u + (un_b - un_inf) * nx;
        qb[2] = p.v + (un_b - un_inf) * ny;
        qb[3] = p.w + (un_b - un_inf) * nz;
        const double rho_inf = p.p / (R * dmax(p.T, 1.0e-6));
        s = p.p / pow(dmax(rho_inf, 1.0e-12), gamma);
    }
    const double rho_b = pow((a_b * a_b) / (gamma * dmax(s, 1.0e-12)), inv_gm1);
    qb[0] = (rho_b * a_b * a_b) / gamma;
    qb[4] = (a_b * a_b) / (gamma * R);
}

}  // namespace cfdg
