/*
 * cfdg.h -- C ABI of the MI355X (gfx950) implicit-iteration library.
 *
 * One context owns one mesh's device-resident data (structure-of-arrays
 * state, face/cell CSR lists, BSR matrix, Krylov work vectors) and one HIP
 * stream.  Every call is synchronous on return, never throws, and returns
 * CFDG_OK (0) or a negative CFDG_ERR_* code; cfdg_last_error() gives the
 * message.  Host buffers are borrowed for the duration of the call only.
 * One host thread per context (the reference runs MPI_THREAD_FUNNELED,
 * apps/solver.cpp:22-43).  Results are deterministic: no floating-point
 * atomics anywhere.
 *
 * Array conventions: "SoA" arrays hold variable v of cell c at
 * a[v * ld + c]; block vectors of the linear system are AoS, x[5 * c + v],
 * exactly as the reference's BSR vectors (implicit_euler.hpp:81-92).
 * Primitive variables are (p, u, v, w, T), conservative (rho, rho u, rho v,
 * rho w, rho E) (state_conversions.hpp:74-129).
 *
 * Reference interfaces replaced (file:line in the reference tree):
 *   cfdg_mesh_*            MeshPart as loaded by import_mesh_hdf5
 *                          (include/cfd/mesh/localmesh.hpp:17-102,
 *                          src/io/solver_mesh/hdf5_reader.cpp:140-461)
 *   cfdg_create            Solver<...> construction: BoundaryManager::initialize,
 *                          allocate_fields, aux connectivity/geometry, gradient
 *                          setup, BSRLinearSystem<5> (solver.hpp:74-150,
 *                          bc_manager.hpp:147-178, bsr_linear_system.hpp:49-80)
 *   cfdg_set_state /       u_slots() conservative SoA state (solver.hpp:291-293)
 *   cfdg_get_state
 *   cfdg_init_uniform      Solver::init_fields (solver.hpp:507-543)
 *   cfdg_eval_residual     Solver::evaluate_residual (solver.hpp:169-247)
 *   cfdg_get_primitives    q slots after BC ghosts (solver.hpp:185-195)
 *   cfdg_get_gradients     gradient planes + limiters (solver.hpp:205-225)
 *   cfdg_get_residual      res slots + spectral radius (residual_kernel.hpp:64-91)
 *   cfdg_residual_norms    Solver::residual_norms (solver.hpp:740-756)
 *   cfdg_compute_dt        Solver::compute_dt (solver.hpp:256-287)
 *   cfdg_assemble_jacobian JacobianKernel::apply (jacobian_kernel.hpp:74-103)
 *   cfdg_get_bsr*          BsrMatrix::{rowPtr, cols, diagIndex, valuesData}
 *                          (include/cfd/linalg/bsr_matrix.hpp:24-62)
 *   cfdg_linear_solve      BSRLinearSystem::solve = SgsPreconditioner::setup +
 *                          BiCGSTAB::solve (bsr_linear_system.hpp:93-97,
 *                          src/linalg/bicgstab.cpp:42-201)
 *   cfdg_sgs_apply         SgsPreconditioner::{setup, apply} (preconditioners.cpp:62-194)
 *   cfdg_spmv              BsrMatrix::apply (bsr_matrix.cpp:10-20)
 *   cfdg_implicit_step     BackwardEuler::advance (implicit_euler.hpp:34-76)
 */
#ifndef CFDG_H
#define CFDG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define CFDG_OK             0
#define CFDG_ERR_ARG       -1  /* invalid argument or size */
#define CFDG_ERR_HIP       -2  /* HIP runtime error (no device, launch failure, ...) */
#define CFDG_ERR_STATE     -3  /* call out of order (e.g. no residual evaluated yet) */
#define CFDG_ERR_SINGULAR  -4  /* singular diagonal block in the SGS setup
                                  (preconditioners.cpp:84-89 is fatal there) */
#define CFDG_ERR_CONFIG    -5  /* unknown / inconsistent configuration value */
#define CFDG_ERR_MESH      -6  /* mesh fails the structural checks */

/* ---- enumerations (names follow the reference's TOML vocabulary) ------------ */
/* flow model: [flow] flow_model (src/solver/compressible/config.cpp:52-62) */
#define CFDG_INVISCID_FLOW 0
#define CFDG_VISCOUS_FLOW  1
/* [numerics] reconstruction */
#define CFDG_FIRST_ORDER   0
#define CFDG_MUSCL         1
/* [numerics] limiter (slope_limiters.hpp) */
#define CFDG_LIMITER_VENKAT     0
#define CFDG_LIMITER_BARTH      1
#define CFDG_LIMITER_VAN_ALBADA 2
#define CFDG_LIMITER_UNLIMITED  3
/* [numerics] gradient; LEAST_SQUARES_FACE selects the node-neighbour LSQ
   operator in the reference (config.cpp:144-147) */
#define CFDG_GREEN_GAUSS_FACE    0
#define CFDG_GREEN_GAUSS_CELL    1
#define CFDG_LEAST_SQUARES_NODE  2
/* boundary condition types, same order as bc::physical::BCType
   (include/cfd/bc/physical/physical_bc.hpp:12-22) */
#define CFDG_BC_SUPERSONIC_INLET       0
#define CFDG_BC_SUPERSONIC_OUTLET      1
#define CFDG_BC_SUBSONIC_INLET         2
#define CFDG_BC_SUBSONIC_OUTLET        3
#define CFDG_BC_SLIP_WALL              4
#define CFDG_BC_NO_SLIP_WALL           5
#define CFDG_BC_NO_SLIP_WALL_HEAT_FLUX 6
#define CFDG_BC_SYMMETRY               7
#define CFDG_BC_FARFIELD               8
/* inflow parameterisation (physical_bc.hpp:24-28) */
#define CFDG_INFLOW_VELOCITY       0
#define CFDG_INFLOW_MACH_ANGLES    1
#define CFDG_INFLOW_MACH_DIRECTION 2
/* preconditioner ([linalg] preconditioner) */
#define CFDG_PRECOND_NONE 0
#define CFDG_PRECOND_SGS  1
/* SGS row ordering on the device:
   MULTICOLOUR - greedy colouring of the face graph, rows of one colour in
                 parallel (a different, parallel Gauss-Seidel ordering);
   SEQUENTIAL  - dependency levels of the natural row order, which performs
                 exactly the reference's sequential sweep (bsr_helpers.hpp:390-435) */
#define CFDG_SGS_MULTICOLOUR 0
#define CFDG_SGS_SEQUENTIAL  1
/* BiCGSTAB status (include/cfd/linalg/types.hpp:36-40) */
#define CFDG_CONVERGED      0
#define CFDG_DIVERGED       1
#define CFDG_MAX_ITERATIONS 2

/* ---- mesh ----------------------------------------------------------------- */
typedef struct cfdg_mesh cfdg_mesh;

/* Synthetic structured-to-unstructured generator (hexahedra, or six Kuhn
   tetrahedra per hexahedron); patches: 0 xmin, 1 xmax, 2 ymin, 3 ymax,
   4 zmin, 5 zmax, 6 ymin for x < split_ymin_x when split_ymin != 0. */
typedef struct {
    int32_t kind;      /* 0 hexahedra, 1 tetrahedra */
    int32_t ni, nj, nk;
    double x0, x1, y0, y1, z0, z1;
    double y_ratio;    /* geometric y growth from y0; 1 = uniform */
    double ramp_x, ramp_deg;          /* bottom shear for x > ramp_x, decaying to the top */
    double bump_h, bump_w, bump_x;    /* bottom Gaussian bump, decaying to the top */
    double jitter;     /* interior node jitter, fraction of the local spacing */
    uint64_t seed;
    int32_t split_ymin;
    double split_ymin_x;
    int32_t reorder;   /* 0 natural order, 1 reverse Cuthill-McKee */
} cfdg_mesh_spec;

/* Borrowed view of a mesh's arrays (layout of MeshPart, localmesh.hpp:17-102). */
typedef struct {
    int32_t n_cells, n_nodes, n_faces, n_inner_faces, n_patches;
    const uint8_t* cell_type;
    const int32_t* cell_nodes_off;   /* n_cells + 1 */
    const int32_t* cell_nodes;
    const double* cell_cx; const double* cell_cy; const double* cell_cz; const double* cell_vol;
    const double* node_x; const double* node_y; const double* node_z;
    const int32_t* face_owner; const int32_t* face_neigh; const int32_t* face_patch;
    const uint8_t* face_type;
    const int32_t* face_nodes_off;   /* n_faces + 1 */
    const int32_t* face_nodes;
    const double* face_cx; const double* face_cy; const double* face_cz;
    const double* face_nx; const double* face_ny; const double* face_nz; const double* face_area;
} cfdg_mesh_view;

int cfdg_mesh_generate(const cfdg_mesh_spec* spec, cfdg_mesh** out);
/* Copies a caller-provided mesh (e.g. read from a container); checks it. */
int cfdg_mesh_from_view(const cfdg_mesh_view* view, cfdg_mesh** out);
int cfdg_mesh_get_view(const cfdg_mesh* mesh, cfdg_mesh_view* view);
int cfdg_mesh_patch_name(const cfdg_mesh* mesh, int32_t patch, char* buf, int32_t buflen);
/* Renumbers cells in reverse Cuthill-McKee order and re-sorts faces
   (src/mesh/reorder.cpp:135-188, 306-576). */
int cfdg_mesh_reorder_rcm(cfdg_mesh* mesh);
void cfdg_mesh_free(cfdg_mesh* mesh);

/* ---- solver context ----------------------------------------------------- */
typedef struct cfdg_ctx cfdg_ctx;

/* SolverConfig subset that drives the residual (config.hpp / config.cpp:28-288) */
typedef struct {
    int32_t flow_model;
    double gamma, gas_constant;
    double prandtl;        /* [flow] prandtl (viscous) */
    double prandtl_turb;   /* turbulent Prandtl number, 0.9 by default (types.hpp:43) */
    int32_t reconstruction;
    int32_t limiter;
    double venkat_k;
    int32_t gradient;
    double cfl;
} cfdg_numerics;

/* One [[boundary_condition]] entry (include/cfd/bc/config.hpp:13-31). */
typedef struct {
    int32_t type;
    int32_t inflow_mode;
    double p, t, tmp_grad;
    double velocity[3];
    double mach, alpha_deg, beta_deg;
    double direction[3];
} cfdg_bc;

/* bcs[p] is the condition of patch p; n_bcs must equal the mesh's patch count. */
int cfdg_create(const cfdg_mesh* mesh, const cfdg_numerics* num, const cfdg_bc* bcs, int32_t n_bcs,
                cfdg_ctx** out);
int cfdg_destroy(cfdg_ctx* ctx);
const char* cfdg_last_error(const cfdg_ctx* ctx);
/* n_total = n_cells + boundary faces (ghost cells follow the cells). */
int cfdg_sizes(const cfdg_ctx* ctx, int32_t* n_cells, int32_t* n_total, int64_t* nnzb);

int cfdg_set_state(cfdg_ctx* ctx, const double* U, int64_t ld);
int cfdg_get_state(cfdg_ctx* ctx, double* U, int64_t ld);
int cfdg_init_uniform(cfdg_ctx* ctx, double rho, const double velocity[3], double p);

int cfdg_eval_residual(cfdg_ctx* ctx);
int cfdg_get_primitives(cfdg_ctx* ctx, double* q, int64_t ld);          /* 5 x n_total */
int cfdg_get_gradients(cfdg_ctx* ctx, double* grad, int64_t ld, double* phi, int64_t ld_phi);
                                                                        /* grad[(3 v + d) ld + c] */
int cfdg_get_residual(cfdg_ctx* ctx, double* R, int64_t ld, double* lam); /* 5 x n_cells, n_cells */
int cfdg_residual_norms(cfdg_ctx* ctx, double l2[5]);
int cfdg_compute_dt(cfdg_ctx* ctx);
int cfdg_get_dt(cfdg_ctx* ctx, double* dt, double* alpha);

int cfdg_assemble_jacobian(cfdg_ctx* ctx);
int cfdg_get_bsr_pattern(cfdg_ctx* ctx, int32_t* row_ptr, int32_t* cols, int32_t* diag_idx);
int cfdg_get_bsr_values(cfdg_ctx* ctx, double* values);          /* nnzb x 25, row-major blocks */
int cfdg_set_bsr_values(cfdg_ctx* ctx, const double* values);

/* [linalg] section (config.cpp:239-285) plus the device options. */
typedef struct {
    int32_t preconditioner;   /* CFDG_PRECOND_* */
    int32_t sgs_order;        /* CFDG_SGS_* */
    double rel_tol, abs_tol;
    int32_t max_iter;
} cfdg_linparams;

typedef struct {
    int32_t status;           /* CFDG_CONVERGED / CFDG_DIVERGED / CFDG_MAX_ITERATIONS */
    int32_t iterations;
    double initial_residual;  /* ||b - A x0|| */
    double final_residual;    /* recurrence residual at exit */
} cfdg_iter_result;

/* Solves A x = b with the current BSR values.  b == NULL uses the implicit
   right-hand side -R; x0 == NULL starts from zero; x (AoS, 5 n_cells) receives
   the solution when non-NULL. */
int cfdg_linear_solve(cfdg_ctx* ctx, const cfdg_linparams* lp, const double* b, const double* x0,
                      double* x, cfdg_iter_result* res);

/* z = M^-1 r for one SGS application with the current BSR values: z = 0, a
   forward and a backward sweep (SgsPreconditioner::apply, preconditioners.cpp:
   171-194; LU of the diagonal blocks as in :62-93).  AoS vectors, 5 n_cells. */
int cfdg_sgs_apply(cfdg_ctx* ctx, int32_t sgs_order, const double* r, double* z);
/* y = A x with the current BSR values (BsrMatrix::apply, bsr_matrix.cpp:10-20). */
int cfdg_spmv(cfdg_ctx* ctx, const double* x, double* y);

typedef struct {
    double l2[5];                 /* residual norms of R(U^n) */
    cfdg_iter_result linear;
} cfdg_step_result;

/* One BackwardEuler::advance: residual, dt, Jacobian, solve (alpha I + J) du = -R,
   U <- U + du. */
int cfdg_implicit_step(cfdg_ctx* ctx, const cfdg_linparams* lp, cfdg_step_result* out);

/* Device time (milliseconds, HIP events on the context stream) of the last
   call of each stage: 0 primitives+ghosts, 1 gradients, 2 limiter,
   3 flux residual, 4 Jacobian assembly, 5 linear solve. */
int cfdg_stage_times(cfdg_ctx* ctx, double ms[6]);
/* Launches the flux-residual kernel reps times back to back on the context
   stream and returns the mean device time per launch in milliseconds. */
int cfdg_time_flux_kernel(cfdg_ctx* ctx, int32_t reps, double* ms_per_launch);

#ifdef __cplusplus
}
#endif

#endif /* CFDG_H */
