/*
 * cfdg_front.h -- the reference solver's host front end over the cfdg
 * library: its configuration files, its mesh container and its run driver.
 *
 * Reference interfaces these entry points replace (one rank, one GPU):
 *   cfdg_config_read_solver    cfd::solver::compressible::parse_solver_config
 *                              (src/solver/compressible/config.cpp:28-288)
 *   cfdg_config_read_boundary  cfd::bc::parse_boundary_config
 *                              (src/bc/config.cpp:141-289)
 *   cfdg_config_to_numerics    the dispatch of run_solver on the parsed config
 *                              (src/solver/compressible/solver.cpp:19-289) onto
 *                              cfdg_numerics / cfdg_linparams
 *   cfdg_container_read        cfd::io::solver_mesh::import_mesh_hdf5
 *                              (src/io/solver_mesh/hdf5_reader.cpp:140-461)
 *   cfdg_container_write       cfd::io::solver_mesh::export_mesh_hdf5
 *                              (src/io/solver_mesh/hdf5_writer.cpp:195-389)
 *   cfdg_run_solver            run_solver + Solver::run (solver.cpp:229-289,
 *                              solver.hpp:152-162, 594-736): init_fields, the
 *                              steady or dual-time loop, progress lines and
 *                              VTU output
 *
 * Messages: where the reference calls mpi::fatal(comm, what) and aborts, these
 * functions return an error code and copy `what` -- the reference's text -- to
 * err (errlen bytes, NUL-terminated; err may be NULL).  The TOML syntax-error
 * descriptions after "config: cannot parse '<path>': " are this reader's own,
 * not toml++'s.
 */
#ifndef CFDG_FRONT_H
#define CFDG_FRONT_H

#include <stdint.h>

#include "cfdg.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CFDG_ERR_IO -8 /* container cannot be read or written */

/* SolverConfig enumerations, in the reference's order
   (include/cfd/solver/compressible/config.hpp:18-95, linalg/config.hpp:8-12,
   numerics/gradient/gradient_manager.hpp:14-19) */
#define CFDG_CFG_EOS_IDEAL_GAS 0
#define CFDG_CFG_FLOW_INVISCID 0
#define CFDG_CFG_FLOW_VISCOUS 1
#define CFDG_CFG_FLUX_HLLC 0
#define CFDG_CFG_SCHEME_FORWARD_EULER 0
#define CFDG_CFG_SCHEME_BACKWARD_EULER 1
#define CFDG_CFG_MODE_STEADY 0
#define CFDG_CFG_MODE_UNSTEADY 1
#define CFDG_CFG_RECON_FIRST_ORDER 0
#define CFDG_CFG_RECON_MUSCL 1
#define CFDG_CFG_RECON_MUSCL_DIRECTIONAL 2
#define CFDG_CFG_LIMITER_NONE 0
#define CFDG_CFG_LIMITER_VENKATAKRISHNAN 1
#define CFDG_CFG_LIMITER_BARTH_JESPERSEN 2
#define CFDG_CFG_LIMITER_VAN_ALBADA 3
#define CFDG_CFG_LIMITER_MINMOD_1D 4
#define CFDG_CFG_LIMITER_VAN_ALBADA_1D 5
#define CFDG_CFG_GRADIENT_GREEN_GAUSS_FACE 0
#define CFDG_CFG_GRADIENT_GREEN_GAUSS_CELL 1
#define CFDG_CFG_GRADIENT_LEAST_SQUARES_CELL_FACE 2
#define CFDG_CFG_GRADIENT_LEAST_SQUARES_CELL_NODE 3
#define CFDG_CFG_SOLVER_BICGSTAB 0
#define CFDG_CFG_PRECOND_NONE 0
#define CFDG_CFG_PRECOND_SGS 1
#define CFDG_CFG_VERBOSITY_SILENT 0
#define CFDG_CFG_VERBOSITY_SUMMARY 1
#define CFDG_CFG_VERBOSITY_VERBOSE 2

/* SolverConfig (config.hpp:97-136), field for field */
typedef struct {
    /* [flow] */
    int32_t eos;
    double gamma, gas_constant, prandtl;
    int32_t flow_model;
    /* [initial] */
    double init_rho, init_p, init_velocity[3];
    /* [numerics] */
    int32_t gradient, flux, reconstruction, limiter;
    double venkat_k;
    /* [turbulence] */
    int32_t turbulence; /* 1: model = "SA" */
    double nu_inf_ratio;
    int32_t max_distance_sweeps;
    double distance_tolerance;
    /* [time] */
    int32_t scheme;
    double cfl;
    int64_t max_iterations;
    double residual_tolerance;
    int32_t time_mode;
    double dt;
    int64_t max_time_steps, bdf_order;
    /* [linalg] (SolverParams defaults when the section is absent) */
    int32_t solver, preconditioner;
    double rel_tol, abs_tol;
    int32_t max_iter, verbosity;
    /* [output] */
    char output_dir[1024];
    int64_t field_interval, residual_interval;
} cfdg_solver_config;

/* Reads a solver configuration file (or, _text, a document held in memory,
   with path used in the messages).  CFDG_ERR_CONFIG with the reference's
   message on any violation, CFDG_ERR_ARG for a directory that does not fit
   output_dir. */
int cfdg_config_read_solver(const char* path, cfdg_solver_config* out, char* err, int32_t errlen);
int cfdg_config_read_solver_text(const char* text, const char* path, cfdg_solver_config* out, char* err,
                                 int32_t errlen);

/* Reads a boundary-condition file against a mesh's patches: bcs[p] receives
   the condition of patch p (n_bcs = the patch count).  warnings (may be NULL)
   receives the reference's log_stat warnings, one per line.  The _text form
   takes the patch names instead of a mesh. */
int cfdg_config_read_boundary(const char* path, const cfdg_mesh* mesh, cfdg_bc* bcs, int32_t n_bcs, char* warnings,
                              int32_t wlen, char* err, int32_t errlen);
int cfdg_config_read_boundary_text(const char* text, const char* path, const char* const* patch_names,
                                   int32_t n_patches, cfdg_bc* bcs, char* warnings, int32_t wlen, char* err,
                                   int32_t errlen);

/* The device's numerics and linear-solver parameters of a parsed config.
   CFDG_ERR_CONFIG (with a message) for what the device path does not run:
   MUSCL_DIRECTIONAL reconstruction and its 1-D limiters.  The SGS order of
   lp is CFDG_SGS_MULTICOLOUR; abs_tol, rel_tol and max_iter are the
   config's. */
int cfdg_config_to_numerics(const cfdg_solver_config* cfg, cfdg_numerics* num, cfdg_linparams* lp, char* err,
                            int32_t errlen);

/* Mesh container of one partition.  Read: the partition of rank 0 of a
   one-rank run (a file written for more ranks is rejected with the
   reference's message), checked like cfdg_mesh_from_view, patch names / CGNS
   types and global ids kept.  Write: the reference's layout with
   rank2part = part2rank = [0] and an empty /comm group.  CFDG_ERR_IO with a
   message when libhdf5 cannot be loaded (CFDG_HDF5_LIB names it; default
   /opt/conda/lib/libhdf5.so.103) or the file does not follow the layout. */
int cfdg_container_read(const char* path, cfdg_mesh** out, char* err, int32_t errlen);
int cfdg_container_write(const char* path, const cfdg_mesh* mesh, char* err, int32_t errlen);

/* The reference's run on the device: the context of (cfg, bcs, mesh),
   init_fields with [initial], then
     steady, BACKWARD_EULER : cfdg_implicit_step per iteration
     steady, FORWARD_EULER  : cfdg_explicit_step per iteration
     unsteady (BDF 1 or 2)  : dual time, BACKWARD_EULER sub-iterations
   with the reference's diagnostics (relative residual at iteration 1, every
   residual_interval and the last; progress lines on stderr in its format;
   boundary integrals per patch when verbose >= 1), solution files
   <output_dir>/iter_<9 digits> every field_interval iterations
   (unsteady_<step> in dual time), final_<last> at the end and blowup on a
   non-finite residual (then 1 is returned).  [turbulence] model = "SA" runs
   the module (point-implicit update in steady runs, frozen in dual time).
   Returns 0 on completion (converged or not), 1 on a non-finite residual,
   or a negative CFDG_ERR_* code with the message in err. */
int cfdg_run_solver(const cfdg_solver_config* cfg, const cfdg_bc* bcs, int32_t n_bcs, const cfdg_mesh* mesh,
                    int32_t verbose, char* err, int32_t errlen);

#ifdef __cplusplus
}
#endif

#endif /* CFDG_FRONT_H */
