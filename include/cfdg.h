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
 *   cfdg_boundary_integrals ResidualKernel::boundary_integrals (residual_kernel.hpp:96-169)
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
 *   cfdg_run_steady        Solver::run_steady (solver.hpp:594-654)
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
/* Patch metadata of MeshPart::patches (localmesh.hpp:86-92): the name and the
   CGNS BC type string; meshes from cfdg_mesh_from_view start as "patch<p>" /
   "UserDefined", generated ones as the generator's names / "UserDefined".
   Either string may be NULL to keep it. */
int cfdg_mesh_patch_cgns_type(const cfdg_mesh* mesh, int32_t patch, char* buf, int32_t buflen);
int cfdg_mesh_set_patch(cfdg_mesh* mesh, int32_t patch, const char* name, const char* cgns_type);
/* Global id of each cell (n_cells values; MeshPart::cell_gid,
   localmesh.hpp:17-102): the cell's index at generation or copy, carried
   through renumbering, so gid[new] = old index after cfdg_mesh_reorder_rcm
   (reorder.cpp:306-576 permutes cell_gid the same way). */
int cfdg_mesh_cell_gid(const cfdg_mesh* mesh, int64_t* gid);
/* Renumbers cells in reverse Cuthill-McKee order and re-sorts faces
   (src/mesh/reorder.cpp:135-188, 306-576). */
int cfdg_mesh_reorder_rcm(cfdg_mesh* mesh);
/* The same renumbering with the order computed on the current HIP device
   (level-synchronous breadth-first passes whose ties are broken as the
   sequential ones break them: pseudo-peripheral start, then children by
   (degree, index) under the first parent in queue order), giving the same
   cell order entry for entry.  CFDG_ERR_HIP on a device failure, the mesh
   then left unchanged. */
int cfdg_mesh_reorder_rcm_device(cfdg_mesh* mesh);
/* Recomputes the cell centroids and volumes and the face centroids, unit
   normals and areas from the nodes and node lists on the device
   (compute_mesh_geometry, src/mesh/geometry.cpp:301-539), with the same
   arithmetic as the host computation of cfdg_mesh_generate, so the values
   are the same bit for bit.  Tetrahedra and hexahedra; faces of 3 to 8
   nodes.  CFDG_ERR_MESH for a degenerate cell or face (volume or area not
   above 1e-15), the mesh then left unchanged. */
int cfdg_mesh_geometry_device(cfdg_mesh* mesh);
/* Builds the cell connectivity the reference's MeshAuxConnectivity holds
   (src/mesh/aux_connectivity.cpp:11-187, 348-445) and the Jacobian's BSR
   pattern (bsr_helpers.hpp:16-59) once for the mesh: the faces of each cell
   in ascending face index, the face neighbours and (node_stencil != 0) the
   vertex neighbours of each cell in ascending order, and the BSR rows
   sort({i} U face neighbours).  on_device != 0 builds them on the current
   HIP device (sorted 64-bit keys), otherwise on the host; both give the same
   arrays entry for entry.  Contexts created from the mesh afterwards use
   them instead of building their own; cfdg_mesh_reorder_rcm drops them.
   CFDG_ERR_MESH when two faces couple the same pair of cells (as context
   creation reports it), CFDG_ERR_HIP on a device failure. */
int cfdg_mesh_connectivity(cfdg_mesh* mesh, int32_t on_device, int32_t node_stencil);
typedef struct {
    int32_t n_cells;
    int32_t has_node_stencil;
    const int32_t* cell_faces_off;      /* n_cells + 1 */
    const int32_t* cell_faces;          /* cell_faces_off[n_cells] */
    const int32_t* cell_cells_face_off; /* n_cells + 1 */
    const int32_t* cell_cells_face;
    const int32_t* cell_cells_node_off; /* n_cells + 1, NULL without the vertex stencil */
    const int32_t* cell_cells_node;
    const int32_t* bsr_row_ptr;         /* n_cells + 1 */
    const int32_t* bsr_cols;            /* bsr_row_ptr[n_cells] */
    const int32_t* bsr_diag;            /* n_cells */
} cfdg_mesh_connectivity_view;
/* Borrowed view of the connectivity built by cfdg_mesh_connectivity (valid
   until the mesh is freed, renumbered or given a new connectivity);
   CFDG_ERR_STATE when none is built. */
int cfdg_mesh_get_connectivity(const cfdg_mesh* mesh, cfdg_mesh_connectivity_view* view);
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
/* Per-patch boundary integrals of the last evaluated residual
   (ResidualKernel::boundary_integrals, residual_kernel.hpp:96-169): mass[p] =
   net mass flux, energy[p] = inviscid + viscous energy flux through patch p.
   force[3 p + k] (no reference counterpart) = sum of the momentum fluxes
   F[1 + k] (inviscid + viscous), i.e. the force of the fluid on the patch.
   Any output may be NULL; n_patches must equal the mesh's patch count. */
int cfdg_boundary_integrals(cfdg_ctx* ctx, double* mass, double* energy, double* force, int32_t n_patches);
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
/* Row groups of the device SGS sweeps (no reference counterpart; they
   define the order that replaces the row loop of bsr_sgs_sweep,
   bsr_helpers.hpp:390-435): fwd_group[i] / bwd_group[i] = the group in which
   the forward / backward sweep updates mesh row i (groups run in ascending
   order, rows of one group at once); n_groups[0..1] = the group counts.
   MULTICOLOUR: forward groups are the colours, backward groups the colours
   in reverse (the backward sweep skips its group 0, which the forward sweep
   has just computed from the same inputs); SEQUENTIAL: the dependency levels
   of the natural forward / backward row order.  Any pointer may be NULL. */
int cfdg_sgs_row_groups(cfdg_ctx* ctx, int32_t sgs_order, int32_t* fwd_group, int32_t* bwd_group,
                        int32_t* n_groups);
/* y = A x with the current BSR values (BsrMatrix::apply, bsr_matrix.cpp:10-20). */
int cfdg_spmv(cfdg_ctx* ctx, const double* x, double* y);

typedef struct {
    double l2[5];                 /* residual norms of R(U^n) */
    cfdg_iter_result linear;
} cfdg_step_result;

/* One BackwardEuler::advance: residual, dt, Jacobian, solve (alpha I + J) du = -R,
   U <- U + du. */
int cfdg_implicit_step(cfdg_ctx* ctx, const cfdg_linparams* lp, cfdg_step_result* out);

/* Steady run: implicit steps until the relative residual
   max_v L2_v(R(U^n)) / max(L2_v(R(U^0)), 1e-300) <= rel_tol, checked at step 1
   and every check_interval steps, or max_steps steps (Solver::run_steady,
   solver.hpp:594-654; residual_norms / relative_residual, :740-765).  A
   non-finite norm stops the run with status CFDG_DIVERGED (the reference
   dumps the state and returns 1). */
typedef struct {
    int64_t max_steps;
    double rel_tol;
    int32_t check_interval;   /* residual_interval, >= 1 */
} cfdg_steady_params;

typedef struct {
    int32_t status;             /* CFDG_CONVERGED / CFDG_DIVERGED / CFDG_MAX_ITERATIONS */
    int64_t steps;
    int64_t linear_iterations;  /* summed over the steps */
    double rel;                 /* last relative residual */
    double l2_first[5], l2_last[5];
    double seconds;             /* wall time of the run */
} cfdg_steady_result;

int cfdg_run_steady(cfdg_ctx* ctx, const cfdg_linparams* lp, const cfdg_steady_params* sp,
                    cfdg_steady_result* out);

/* ---- Time integration other than the implicit pseudo-time step --------------
 *
 * cfdg_explicit_step   ForwardEuler::advance (forward_euler.hpp:24-30): the
 *                      residual, the explicit local step alpha = CFL / lambda
 *                      (solver.hpp:278-285, CFL of the numerics), then
 *                      U <- U - alpha R (update_ops.hpp:28-44).  out->l2 = the
 *                      norms of that residual; out->linear is zero.  Not with
 *                      the turbulence module or in dual time (CFDG_ERR_CONFIG /
 *                      CFDG_ERR_STATE).
 * cfdg_set_unsteady    dual time stepping (solver.hpp:355-380, 494-528): BDF
 *                      order 1 or 2 with physical step dt, or 0 for steady; the
 *                      time levels U^n = U^{n-1} = the current state and the
 *                      physical step count restarts.  Every residual then adds
 *                      V (c0 U + c1 U^n [+ c2 U^{n-1}]) (solver.hpp:546-575) and
 *                      the implicit diagonal shift adds V c0 / dt, c0 = 1.5
 *                      (BDF2) or 1 (BDF1, and the first physical step).
 * cfdg_run_unsteady    Solver::run_dual_time (solver.hpp:662-736): `steps`
 *                      physical steps, each of implicit sub-iterations until
 *                      the relative residual (checked at sub-iteration 1, every
 *                      5th and the last) reaches sub_tol or max_subiter runs
 *                      out, then U^{n-1} = U^n, U^n = U.  subiters[steps], if
 *                      given, receives the sub-iterations of each step.  The
 *                      physical step count carries over between calls.
 */
int cfdg_explicit_step(cfdg_ctx* ctx, cfdg_step_result* out);
/* ---- Solution output ---------------------------------------------------------
 *
 * cfdg_write_solution_vtu  Solver::write_fields (solver.hpp:856-898) with
 *                          io::vtk::write_solution_vtu (vtu.cpp:321-475) for
 *                          one rank: <outdir>/<stem>_00000.vtu (the mesh, a
 *                          rank array and the cell fields rho, u, v, w,
 *                          pressure, mach of the current state, formed on the
 *                          device; with the turbulence module also nu_tilde and
 *                          mu_t of the last residual evaluation, zero before
 *                          the first, as the module appends them,
 *                          spalart_allmaras.hpp:445-448) and its master
 *                          <outdir>/<stem>.pvtu, in the reference's text
 *                          formats; outdir is created if needed.  mesh = the
 *                          mesh the context was created on.
 */
int cfdg_write_solution_vtu(cfdg_ctx* ctx, const cfdg_mesh* mesh, const char* outdir, const char* stem);
int cfdg_set_unsteady(cfdg_ctx* ctx, double dt, int32_t order);
int cfdg_run_unsteady(cfdg_ctx* ctx, const cfdg_linparams* lp, int64_t steps, int64_t max_subiter, double sub_tol,
                      int64_t* subiters);

/* ---- precision forms ------------------------------------------------------
   No counterpart in the reference, whose linear algebra is FP64 throughout
   (BsrMatrix values are double, bsr_matrix.hpp:24-62).  The default form is
   FP64 blocks + FP64 Krylov, i.e. the reference's arithmetic.  With
   krylov = FP32 the Jacobian blocks are stored in FP32 or BF16 (the SGS
   diagonal LU is factorised in FP64 from the stored blocks and kept in FP32),
   and cfdg_linear_solve becomes an FP64 defect correction: r = b - A x in FP64
   arithmetic over the stored blocks, an FP32 BiCGSTAB + FP32 SGS solve of
   A e = r to max(inner_rel_tol ||r||, 0.9 tol) with tol the caller's
   tolerance below, x += e, repeated until ||r|| meets the caller's
   max(abs_tol, rel_tol ||b||) or max_iter inner iterations / max_outer passes
   are spent.  iterations counts inner iterations; final_residual is the FP64
   ||b - A x||.  Dot products accumulate in FP64 in every form.  The flux
   residual, gradients, norms and the state update stay FP64. */
#define CFDG_PREC_FP64 0
#define CFDG_PREC_FP32 1
#define CFDG_PREC_BF16 2

typedef struct {
    int32_t block_storage;  /* CFDG_PREC_FP64 | CFDG_PREC_FP32 | CFDG_PREC_BF16 */
    int32_t krylov;         /* CFDG_PREC_FP64 (needs FP64 blocks) | CFDG_PREC_FP32 */
    int32_t max_outer;      /* defect-correction passes (>= 1) */
    int32_t reserved;       /* 0 */
    double inner_rel_tol;   /* each inner solve stops at max(inner_rel_tol ||r||, 0.9 tol) (0 < inner_rel_tol < 1) */
} cfdg_precision;

/* Selects the form.  When the block storage changes it is reallocated, and
   cfdg_linear_solve, cfdg_sgs_apply and cfdg_spmv return CFDG_ERR_STATE until
   cfdg_assemble_jacobian or cfdg_set_bsr_values fills it (as they do on a
   new context); with the same storage the blocks are kept.  cfdg_get_bsr_values /
   cfdg_set_bsr_values convert between the storage and double; cfdg_spmv
   multiplies the stored blocks in FP64; cfdg_sgs_apply runs the form's SGS. */
int cfdg_set_precision(cfdg_ctx* ctx, const cfdg_precision* prec);
int cfdg_get_precision(cfdg_ctx* ctx, cfdg_precision* prec);

/* ---- turbulence module ----------------------------------------------------
   Replaces the reference's Spalart-Allmaras physics module (SpalartAllmaras,
   spalart_allmaras.hpp:61-544; config [turbulence] model / nu_inf_ratio /
   max_distance_sweeps / distance_tolerance, config.cpp:155-189), which does
   not compile there.  One transported variable rho nu_tilde per cell.  Each
   residual evaluation adds, in the reference's order (solver.hpp:195-241):
   zone ghosts (no-slip wall: -inner; inlet / farfield: rho_g nu_inf;
   outlet, slip wall, symmetry: inner), density / nu_tilde / eddy viscosity
   mu_t = rho nu_tilde fv1 (ghosts 0), the nu_tilde gradient with the
   configured method, mu_t in the mean viscous fluxes and Jacobian blocks, the
   face mass flux F[0] recorded by the flux sweep, upwind convection and
   skew-corrected diffusion (interior faces add their spectral radius to
   lambda) and the production / destruction / cb2 sources.  Four corrections
   of the reference's code: the diffusive flux leaves the owner as
   -coef (grad nu . n) A, the cb2 term is a source, the freestream override
   of a boundary face with inflowing mass applies to inflow / outflow zones
   only, and fw = g ((1 + cw3^6) / (g^6 + cw3^6))^(1/6) as the header's
   formula states (its code drops the cw3^6 of the denominator).
   Update after each implicit step: CFDG_TURB_FROZEN copies the block
   unchanged as the reference's implicit Euler does (implicit_euler.hpp:64-72);
   CFDG_TURB_POINT_IMPLICIT sets rho nu_tilde -= res / (V / dt + D) with D the
   diagonal of the module's upwind convection, diffusion and destruction
   linearisation.  Both end with the positivity clamp rho nu_tilde >= 0
   (spalart_allmaras.hpp:435-442). */
#define CFDG_TURB_NONE 0
#define CFDG_TURB_SA   1
#define CFDG_TURB_FROZEN         0
#define CFDG_TURB_POINT_IMPLICIT 1

typedef struct {
    int32_t model;               /* CFDG_TURB_NONE | CFDG_TURB_SA (needs the viscous flow model) */
    int32_t update;              /* CFDG_TURB_FROZEN | CFDG_TURB_POINT_IMPLICIT */
    double nu_inf_ratio;         /* freestream nu_tilde / nu (reference default 3) */
    int32_t max_distance_sweeps; /* wall-distance relaxation sweeps (500) */
    int32_t reserved;            /* 0 */
    double distance_tolerance;   /* wall-distance relative tolerance (1e-8) */
    double rho_inf, p_inf;       /* freestream density and pressure ([initial]) */
} cfdg_turbulence;

/* SpalartAllmaras::initialize + init_state (spalart_allmaras.hpp:157-193):
   zone kinds from the patches' conditions, wall distance from the no-slip
   walls over the vertex stencil, nu_inf = ratio mu(T_inf) / rho_inf, every
   cell at rho_inf nu_inf.  mesh must be the mesh of cfdg_create. */
int cfdg_set_turbulence(cfdg_ctx* ctx, const cfdg_mesh* mesh, const cfdg_turbulence* turb);
/* rho nu_tilde of every cell (mesh order) */
int cfdg_set_turbulence_state(cfdg_ctx* ctx, const double* rho_nu);
/* Prescribed eddy viscosity mu_t of every cell and boundary ghost (n_total
   values, mesh order, ghost of boundary face f at n_cells + f - n_inner), for
   a host-side turbulence model: the mut argument of ResidualKernel::apply
   (residual_kernel.hpp:64-91; the viscous face flux reads both sides, ghosts
   included, :153-154, 231-232, 284-285) and of JacobianKernel::apply
   (jacobian_kernel.hpp:74-90; zero on the ghost side, :252-254), i.e. what
   Solver::mut_ptr() hands them (solver.hpp:230, 253, 942-945).  The field
   stays until changed; mut == NULL clears it.  Needs the viscous flow model;
   CFDG_ERR_STATE while the turbulence module (which forms its own mu_t) is
   on; cfdg_set_turbulence with a model clears it. */
int cfdg_set_eddy_viscosity(cfdg_ctx* ctx, const double* mut, int32_t n_total);
/* Fields of the module after the last residual evaluation, mesh order, each
   argument optional: rho_nu, nu, mut, dist, res (n_cells), grad
   (3 x n_cells), mdot (n_faces, face mass flux), diag (point-implicit
   diagonal, n_cells). */
int cfdg_get_turbulence(cfdg_ctx* ctx, double* rho_nu, double* nu, double* mut, double* dist, double* res,
                        double* grad, double* mdot, double* diag);

/* ---- learned warm start ---------------------------------------------------
   No counterpart in the reference (its BiCGSTAB always starts from x = 0,
   implicit_euler.hpp:81-92; the solver API takes the guess in x,
   bicgstab.cpp:56-60).  A two-layer edge-convolution network over the face
   graph (hidden width 64) predicts the update du of each implicit step from
   per-cell features (state, residual over spectral radius -R / lambda, volume, local time
   step, freestream Mach) and face geometry.  Node products run in BF16 on
   MFMA with FP32 accumulation; training (MSE on normalised updates, Adam)
   uses the library's own forward and backward kernels.  The prediction is
   the initial guess of the linear solve; when its initial residual
   ||b - A x0|| exceeds ||b|| the solve restarts from zero (fallback), and the
   linear tolerance is unchanged.  Parameter layout: see cfdg_gnn_get_params. */
typedef struct cfdg_gnn cfdg_gnn;

/* Glorot-uniform weights from the seed, zero biases. */
int cfdg_gnn_create(uint64_t seed, cfdg_gnn** out);
int cfdg_gnn_destroy(cfdg_gnn* gnn);
const char* cfdg_gnn_last_error(const cfdg_gnn* gnn);
int cfdg_gnn_num_params(int32_t* n);
/* Flat FP32 parameters in the order [WP1; WQ1] (128 x 32, columns >= 24
   zero), G1 (64 x 4), b1 (64), [WP2; WQ2] (128 x 64), G2 (64 x 4), b2 (64),
   Wo (5 x 64), bo (5); all row-major.  set_params resets the Adam state. */
int cfdg_gnn_get_params(cfdg_gnn* gnn, float* params, int32_t n);
int cfdg_gnn_set_params(cfdg_gnn* gnn, const float* params, int32_t n);

/* Freestream of the feature normalisation (cfdg_init_uniform sets it too). */
int cfdg_set_freestream(cfdg_ctx* ctx, double rho, const double velocity[3], double p);
/* on != 0: every implicit step keeps its features; cfdg_gnn_record then adds
   the last step's (features, du) pair as one training sample (the context's
   graph is copied once per context). */
int cfdg_gnn_set_recording(cfdg_ctx* ctx, int32_t on);
int cfdg_gnn_record(cfdg_gnn* gnn, cfdg_ctx* ctx);
int cfdg_gnn_num_samples(cfdg_gnn* gnn, int32_t* n);
/* Adam (beta 0.9 / 0.999, eps 1e-8) over all samples, epochs passes in a
   seeded order, one step per sample; mean_loss = mean loss of the last epoch. */
int cfdg_gnn_train(cfdg_gnn* gnn, int32_t epochs, float lr, uint64_t seed, double* mean_loss);
/* Loss and full parameter gradient of one sample, parameters unchanged. */
int cfdg_gnn_loss_grad(cfdg_gnn* gnn, int32_t sample, double* loss, float* grad, int32_t n);
/* A recorded sample: sizes (any array may be NULL), features n x 32 (the
   stored BF16 values), targets n x 5, graph (device order) with per-edge
   geometry e x 4. */
int cfdg_gnn_get_sample(cfdg_gnn* gnn, int32_t sample, int32_t* n_nodes, int32_t* n_edges, float* x, float* t,
                        int32_t* adj_off, int32_t* adj, int32_t* rev, float* eg);
/* Predicted du (AoS, mesh order) for the current state; needs
   cfdg_eval_residual and cfdg_compute_dt. */
int cfdg_gnn_predict(cfdg_gnn* gnn, cfdg_ctx* ctx, double* du);
/* Initial guess of cfdg_implicit_step / cfdg_linear_solve(b = NULL, x0 =
   NULL): the network's prediction (gnn != NULL) or zero (gnn == NULL).  The
   network must outlive its use by the context. */
int cfdg_set_warm_start(cfdg_ctx* ctx, cfdg_gnn* gnn);
/* Solves since cfdg_set_warm_start that kept / rejected the prediction. */
int cfdg_warm_start_stats(cfdg_ctx* ctx, int64_t* used, int64_t* rejected);

/* Device time (milliseconds, HIP events on the context stream) of the last
   call of each stage: 0 primitives+ghosts, 1 gradients, 2 limiter,
   3 flux residual, 4 Jacobian assembly, 5 linear solve. */
int cfdg_stage_times(cfdg_ctx* ctx, double ms[6]);
/* Launches the flux-residual kernel reps times back to back on the context
   stream and returns the mean device time per launch in milliseconds. */
int cfdg_time_flux_kernel(cfdg_ctx* ctx, int32_t reps, double* ms_per_launch);
/* The form of the gradient + limiter kernel this context runs (no reference
   counterpart; for tests and tuning).  form: CFDG_GRAD_GATHER (one thread per
   cell gathering its stencil from the cell planes: Green-Gauss gradients),
   CFDG_GRAD_LDS (node stencils staged in LDS, at most 640 slots per tile) or
   CFDG_GRAD_LDS_WIDE (1088 slots).  n_gradient_tiles: workgroups of the LDS
   forms, more than the n_flux_tiles when tiles are split into sub-ranges
   whose stencils fit.  Any output may be NULL. */
#define CFDG_GRAD_GATHER   0
#define CFDG_GRAD_LDS      1
#define CFDG_GRAD_LDS_WIDE 2
int cfdg_gradient_form(const cfdg_ctx* ctx, int32_t* form, int32_t* n_gradient_tiles, int32_t* n_flux_tiles);

/* ---- allocation-free entry point -------------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
   memory: code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc.  Computes what cfdg_residual_norms computes
   (Solver::residual_norms, solver.hpp:740-756: l2[v] = sqrt(sum_c R_v^2 /
   n_cells)) for a residual the caller holds in device memory, variable v of
   cell c at res[v * ld + c], and writes the five norms to l2 (device memory).
   It shares the kernels of cfdg_residual_norms; their partial sums go to a
   static device array of the library (1024 x 5 doubles = 40 KiB per GPU,
   held until the process ends) instead of a context buffer, so the call
   allocates nothing.
   - Works on the current device, stream-ordered on `stream` (a hipStream_t;
     NULL = the null stream), and returns without synchronising: read l2 after
     the stream has reached it.
   - Every call on one GPU shares that GPU's scratch array; calls that overlap
     on one GPU (different streams or threads, or graph replays in flight
     together) race and must be serialised by the caller.
   - The first call on each device looks up the array's device address; a
     caller that captures the call into a graph makes one call on that device
     before capturing.
   - n_cells = 0 gives zeros.  Sizes: 0 <= n_cells <= CFDG_NOALLOC_MAX_CELLS
     (larger -> CFDG_ERR_CAPACITY, nothing launched); n_cells <= ld <=
     INT32_MAX; NULL res or l2 -> CFDG_ERR_ARG. */
#define CFDG_ERR_CAPACITY      -7          /* input larger than a fixed capacity */
#define CFDG_NOALLOC_MAX_CELLS 1073741824  /* 2^30 cells */
int cfdg_residual_norms_noalloc(const double* res, int64_t ld, int32_t n_cells, double* l2, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CFDG_H */
