/*
 * oracle.h -- CPU restatement of the reference solver's single-rank implicit
 * iteration.  TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the
 * CPU-baseline leg of the benchmark link it to compare against; the product never
 * does.
 *
 * The structs below have exactly the layout of their cfdg.h counterparts so a
 * test can hand the same argument blocks to both libraries.
 */
#ifndef CFD_ORACLE_H
#define CFD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t n_cells, n_nodes, n_faces, n_inner_faces, n_patches;
    const uint8_t* cell_type;
    const int32_t* cell_nodes_off;
    const int32_t* cell_nodes;
    const double* cell_cx; const double* cell_cy; const double* cell_cz; const double* cell_vol;
    const double* node_x; const double* node_y; const double* node_z;
    const int32_t* face_owner; const int32_t* face_neigh; const int32_t* face_patch;
    const uint8_t* face_type;
    const int32_t* face_nodes_off;
    const int32_t* face_nodes;
    const double* face_cx; const double* face_cy; const double* face_cz;
    const double* face_nx; const double* face_ny; const double* face_nz; const double* face_area;
} oracle_mesh_view;

typedef struct {
    int32_t flow_model;
    double gamma, gas_constant, prandtl, prandtl_turb;
    int32_t reconstruction, limiter;
    double venkat_k;
    int32_t gradient;
    double cfl;
} oracle_numerics;

typedef struct {
    int32_t type, inflow_mode;
    double p, t, tmp_grad;
    double velocity[3];
    double mach, alpha_deg, beta_deg;
    double direction[3];
} oracle_bc;

typedef struct {
    int32_t preconditioner, sgs_order;
    double rel_tol, abs_tol;
    int32_t max_iter;
} oracle_linparams;

typedef struct {
    int32_t status, iterations;
    double initial_residual, final_residual;
} oracle_iter_result;

typedef struct {
    double l2[5];
    oracle_iter_result linear;
} oracle_step_result;

/* Turbulence module (layout of cfdg_turbulence) */
typedef struct {
    int32_t model, update;
    double nu_inf_ratio;
    int32_t max_distance_sweeps, reserved;
    double distance_tolerance;
    double rho_inf, p_inf;
} oracle_turbulence;

typedef struct oracle_ctx oracle_ctx;

int oracle_create(const oracle_mesh_view* mesh, const oracle_numerics* num, const oracle_bc* bcs,
                  int32_t n_bcs, oracle_ctx** out);
int oracle_destroy(oracle_ctx* ctx);
const char* oracle_last_error(const oracle_ctx* ctx);
int oracle_sizes(const oracle_ctx* ctx, int32_t* n_cells, int32_t* n_total, int64_t* nnzb);

int oracle_set_state(oracle_ctx* ctx, const double* U, int64_t ld);
int oracle_get_state(oracle_ctx* ctx, double* U, int64_t ld);
int oracle_init_uniform(oracle_ctx* ctx, double rho, const double velocity[3], double p);

int oracle_eval_residual(oracle_ctx* ctx);
int oracle_get_primitives(oracle_ctx* ctx, double* q, int64_t ld);
int oracle_get_gradients(oracle_ctx* ctx, double* grad, int64_t ld, double* phi, int64_t ld_phi);
int oracle_get_residual(oracle_ctx* ctx, double* R, int64_t ld, double* lam);
/* S[v ld + c] = sum over the faces of cell c of |F_v| (the magnitude scale of
   the residual sum, used to state per-cell tolerances) */
int oracle_get_flux_scale(oracle_ctx* ctx, double* S, int64_t ld);
int oracle_residual_norms(oracle_ctx* ctx, double l2[5]);
int oracle_compute_dt(oracle_ctx* ctx);
int oracle_get_dt(oracle_ctx* ctx, double* dt, double* alpha);

int oracle_assemble_jacobian(oracle_ctx* ctx);
int oracle_get_bsr_pattern(oracle_ctx* ctx, int32_t* row_ptr, int32_t* cols, int32_t* diag_idx);
int oracle_get_bsr_values(oracle_ctx* ctx, double* values);
int oracle_set_bsr_values(oracle_ctx* ctx, const double* values);

int oracle_linear_solve(oracle_ctx* ctx, const oracle_linparams* lp, const double* b, const double* x0,
                        double* x, oracle_iter_result* res);
int oracle_implicit_step(oracle_ctx* ctx, const oracle_linparams* lp, oracle_step_result* out);
/* forward Euler step with the explicit dt (forward_euler.hpp:24-30); l2 of the residual before the update */
int oracle_explicit_step(oracle_ctx* ctx, double l2[5]);
/* dual time stepping: order 0 (steady), 1 or 2; U^n = U^{n-1} = the current state */
int oracle_set_unsteady(oracle_ctx* ctx, double dt, int32_t order);
/* Solver::run_dual_time (solver.hpp:662-736); subiters[steps] receives the sub-iterations per physical step */
int oracle_run_unsteady(oracle_ctx* ctx, const oracle_linparams* lp, int64_t steps, int64_t max_subiter,
                        double sub_tol, int64_t* subiters);

/* Greedy colouring of the face graph in row order (colour = smallest colour
   not used by an already-coloured lower neighbour) and the dependency levels
   of the natural forward / backward sweeps; out arrays have n_cells entries.
   Returns the number of colours / levels through the counters. */
int oracle_row_groups(oracle_ctx* ctx, int32_t* colour, int32_t* n_colours, int32_t* fwd_level,
                      int32_t* n_fwd, int32_t* bwd_level, int32_t* n_bwd);

/* One SGS preconditioner application z = M^-1 r with the current matrix
   (sgs_order selects sequential or multicolour row order). */
int oracle_sgs_apply(oracle_ctx* ctx, int32_t sgs_order, const double* r, double* z);
/* y = A x (AoS block vectors) */
int oracle_spmv(oracle_ctx* ctx, const double* x, double* y);
int oracle_boundary_integrals(oracle_ctx* ctx, double* mass, double* energy, double* force, int32_t n_patches);
int oracle_wall_distance(oracle_ctx* ctx, int32_t stencil, int32_t max_sweeps, double rel_tol, double* dist);

/* Spalart-Allmaras module: set up (zone kinds, wall distance over the vertex
   stencil, freestream nu_tilde = ratio mu(T_inf) / rho_inf, state = rho_inf
   nu_inf), state, and the per-evaluation fields of the last residual (n_cells
   each; grad 3 x n_cells; mdot n_faces; diag the point-implicit diagonal). */
int oracle_set_turbulence(oracle_ctx* ctx, const oracle_turbulence* t);
/* Prescribed eddy viscosity mu_t of every cell and boundary ghost (n_total,
   ghosts after the cells) for the mean viscous flux and Jacobian; NULL clears
   it.  Not with the turbulence module. */
int oracle_set_eddy_viscosity(oracle_ctx* ctx, const double* mut, int32_t n_total);
int oracle_set_turbulence_state(oracle_ctx* ctx, const double* rho_nu);
int oracle_get_turbulence(oracle_ctx* ctx, double* rho_nu, double* nu, double* mut, double* dist, double* res,
                          double* grad, double* mdot, double* diag);
/* One cell's source terms: out = P, D, cb2 term, fv1, fv2, Sbar, r, fw */
int oracle_sa_point(double rho, double nu, double mu_lam, double d, double S, double grad2, double out[8]);
/* Model constants: cb1, cb2, sigma, kappa, cw1, cw2, cw3, cv1 */
double oracle_sa_constant(int32_t which);

#ifdef __cplusplus
}
#endif

#endif
