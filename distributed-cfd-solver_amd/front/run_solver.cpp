// cfdg_run_solver: the reference's run_solver + Solver::run on the device
// (src/solver/compressible/solver.cpp:19-289, include/cfd/solver/compressible/
// solver.hpp:152-162, 594-736, 767-779, 826-898).  Progress lines go to
// stderr in the reference's formats (its mpi::log_* print the formatted
// line and a newline on stderr); the linear solver's summary line goes to
// stdout as BiCGSTAB::solve prints it (bicgstab.cpp:189-199).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "cfdg.h"
#include "cfdg_front.h"

namespace {

void logf(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vfprintf(stderr, fmt, ap);
    va_end(ap);
    std::fputc('\n', stderr);
}

void put_msg(char* buf, int32_t len, const std::string& s) {
    if (buf && len > 0) std::snprintf(buf, static_cast<std::size_t>(len), "%s", s.c_str());
}

// make_stem (solver.hpp:850-854)
std::string stem(const char* tag, long long iter) {
    char b[64];
    std::snprintf(b, sizeof b, "%s_%09lld", tag, iter);
    return b;
}

// relative_residual (solver.hpp:758-765) with kResidualNormFloor = 1e-300
double relative(const double* l2, const double* n0) {
    double r = 0.0;
    for (int v = 0; v < 5; ++v) r = std::max(r, l2[v] / std::max(n0[v], 1.0e-300));
    return r;
}

const char* limiter_name(const cfdg_solver_config& c) {
    if (c.reconstruction == CFDG_CFG_RECON_FIRST_ORDER) return "NONE";
    switch (c.limiter) {
        case CFDG_CFG_LIMITER_VENKATAKRISHNAN: return "VENKATAKRISHNAN";
        case CFDG_CFG_LIMITER_BARTH_JESPERSEN: return "BARTH_JESPERSEN";
        case CFDG_CFG_LIMITER_VAN_ALBADA: return "VAN_ALBADA";
        default: return "?";
    }
}

class Run {
public:
    Run(const cfdg_solver_config& cfg, const cfdg_mesh* mesh, int verbose) : cfg_(cfg), mesh_(mesh), verbose_(verbose) {}
    ~Run() {
        if (ctx_) cfdg_destroy(ctx_);
    }

    int start(const cfdg_bc* bcs, int32_t n_bcs, char* err, int32_t errlen) {
        int rc = cfdg_config_to_numerics(&cfg_, &num_, &lp_, err, errlen);
        if (rc != CFDG_OK) return rc;
        if (cfg_.time_mode == CFDG_CFG_MODE_UNSTEADY && cfg_.bdf_order != 1 && cfg_.bdf_order != 2) {
            put_msg(err, errlen, "dispatch: invalid bdf_order (available: 1, 2)");
            return CFDG_ERR_CONFIG;
        }
        if (cfg_.scheme == CFDG_CFG_SCHEME_FORWARD_EULER &&
            (cfg_.time_mode == CFDG_CFG_MODE_UNSTEADY || cfg_.turbulence)) {
            put_msg(err, errlen,
                    "dispatch: FORWARD_EULER runs on the device in steady mode without turbulence only");
            return CFDG_ERR_CONFIG;
        }
        rc = cfdg_create(mesh_, &num_, bcs, n_bcs, &ctx_);
        if (rc != CFDG_OK) return fail(rc, err, errlen);
        rc = cfdg_init_uniform(ctx_, cfg_.init_rho, cfg_.init_velocity, cfg_.init_p);  // init_fields
        if (rc != CFDG_OK) return fail(rc, err, errlen);
        if (cfg_.turbulence) {
            cfdg_turbulence tb{};
            tb.model = CFDG_TURB_SA;
            tb.update = cfg_.time_mode == CFDG_CFG_MODE_UNSTEADY ? CFDG_TURB_FROZEN : CFDG_TURB_POINT_IMPLICIT;
            tb.nu_inf_ratio = cfg_.nu_inf_ratio;
            tb.max_distance_sweeps = cfg_.max_distance_sweeps;
            tb.distance_tolerance = cfg_.distance_tolerance;
            tb.rho_inf = cfg_.init_rho;
            tb.p_inf = cfg_.init_p;
            rc = cfdg_set_turbulence(ctx_, mesh_, &tb);
            if (rc != CFDG_OK) return fail(rc, err, errlen);
        }
        int32_t nc = 0, nt = 0;
        int64_t nnzb = 0;
        cfdg_sizes(ctx_, &nc, &nt, &nnzb);
        nc_ = nc;
        physics_ = cfg_.flow_model == CFDG_CFG_FLOW_VISCOUS ? "VISCOUS_FLOW" : "INVISCID_FLOW";
        if (cfg_.turbulence) physics_ += "+SA";
        return CFDG_OK;
    }

    int steady(char* err, int32_t errlen) {
        const auto t0 = std::chrono::steady_clock::now();
        const bool implicit = cfg_.scheme == CFDG_CFG_SCHEME_BACKWARD_EULER;
        logf("solver: physics=%s flux=%s recon=%s limiter=%s scheme=%s cfl=%.3f max_iter=%lld", physics_.c_str(),
             "HLLC", recon_name(), limiter_name(cfg_), implicit ? "BACKWARD_EULER" : "FORWARD_EULER", cfg_.cfl,
             static_cast<long long>(cfg_.max_iterations));
        double norm0[5] = {0, 0, 0, 0, 0};
        bool have0 = false;
        long long last = 0;
        for (long long it = 1; it <= cfg_.max_iterations; ++it) {
            last = it;
            cfdg_step_result st{};
            const int rc = implicit ? cfdg_implicit_step(ctx_, &lp_, &st) : cfdg_explicit_step(ctx_, &st);
            if (rc != CFDG_OK) return fail(rc, err, errlen);
            if (implicit && cfg_.verbosity >= CFDG_CFG_VERBOSITY_SUMMARY) linear_summary(st.linear);
            const bool diagnose = it == 1 || it % cfg_.residual_interval == 0 || it == cfg_.max_iterations;
            bool converged = false;
            double rel = 1.0;
            if (diagnose) {
                if (!std::isfinite(st.l2[0])) {
                    std::fprintf(stderr, "WARNING: [r0] solver: non-finite residual detected, dumping state\n");
                    write("blowup");
                    return 1;
                }
                if (!have0) {
                    std::copy(st.l2, st.l2 + 5, norm0);
                    have0 = true;
                } else {
                    rel = relative(st.l2, norm0);
                }
                progress(it, seconds_since(t0), st.l2, rel);
                if (verbose_ >= 1 && boundary_integrals() != CFDG_OK) return fail(CFDG_ERR_STATE, err, errlen);
                converged = rel <= cfg_.residual_tolerance;
            }
            if (cfg_.field_interval > 0 && it % cfg_.field_interval == 0) write(stem("iter", it));
            if (converged) {
                logf("solver: converged at iteration %lld (rel=%.3e)", it, rel);
                break;
            }
        }
        write(stem("final", last));
        logf("solver: done in %lld iterations, wall time %.3f s", last, seconds_since(t0));
        return write_rc_ == CFDG_OK ? 0 : fail(write_rc_, err, errlen);
    }

    int dual_time(char* err, int32_t errlen) {
        const auto t0 = std::chrono::steady_clock::now();
        const int order = static_cast<int>(cfg_.bdf_order);
        logf("solver (unsteady): physics=%s flux=%s recon=%s limiter=%s scheme=%s dt=%.3e total_steps=%lld "
             "max_subiter=%lld",
             physics_.c_str(), "HLLC", recon_name(), limiter_name(cfg_), "BACKWARD_EULER", cfg_.dt,
             static_cast<long long>(cfg_.max_time_steps), static_cast<long long>(cfg_.max_iterations));
        int rc = cfdg_set_unsteady(ctx_, cfg_.dt, order);
        if (rc != CFDG_OK) return fail(rc, err, errlen);
        double t = 0.0;
        for (long long step = 1; step <= cfg_.max_time_steps; ++step) {
            t += cfg_.dt;
            int64_t sub = 0;
            rc = cfdg_run_unsteady(ctx_, &lp_, 1, cfg_.max_iterations, cfg_.residual_tolerance, &sub);
            if (rc == CFDG_ERR_STATE && std::string(cfdg_last_error(ctx_)).find("non-finite") != std::string::npos) {
                std::fprintf(stderr, "WARNING: [r0] solver: non-finite residual detected, dumping state\n");
                write("blowup");
                return 1;
            }
            if (rc != CFDG_OK) return fail(rc, err, errlen);
            if (cfg_.field_interval > 0 && step % cfg_.field_interval == 0) write(stem("unsteady", step));
            logf("Time step %lld/%lld: t = %.5e s | subiters: %lld/%lld (%s)", step,
                 static_cast<long long>(cfg_.max_time_steps), t, static_cast<long long>(sub),
                 static_cast<long long>(cfg_.max_iterations), sub < cfg_.max_iterations ? "converged" : "limit reached");
        }
        write(stem("final", cfg_.max_time_steps));
        logf("solver: unsteady calculation finished in %lld steps, wall time %.3f s",
             static_cast<long long>(cfg_.max_time_steps), seconds_since(t0));
        return write_rc_ == CFDG_OK ? 0 : fail(write_rc_, err, errlen);
    }

private:
    const cfdg_solver_config& cfg_;
    const cfdg_mesh* mesh_;
    int verbose_;
    cfdg_ctx* ctx_ = nullptr;
    cfdg_numerics num_{};
    cfdg_linparams lp_{};
    std::string physics_;
    int32_t nc_ = 0;
    int write_rc_ = CFDG_OK;

    static double seconds_since(std::chrono::steady_clock::time_point t0) {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    const char* recon_name() const { return cfg_.reconstruction == CFDG_CFG_RECON_MUSCL ? "MUSCL" : "FIRST_ORDER"; }

    int fail(int rc, char* err, int32_t errlen) const {
        put_msg(err, errlen, ctx_ ? std::string(cfdg_last_error(ctx_)) : std::string("cfdg_create failed"));
        return rc;
    }

    // write_fields (solver.hpp:856-898); the first failure is kept and reported at the end
    void write(const std::string& s) {
        const int rc = cfdg_write_solution_vtu(ctx_, mesh_, cfg_.output_dir, s.c_str());
        if (rc != CFDG_OK && write_rc_ == CFDG_OK) write_rc_ = rc;
    }

    // log_progress (solver.hpp:767-779); dt of the step just taken
    void progress(long long it, double wall, const double* l2, double rel) {
        const double mom = std::sqrt(l2[1] * l2[1] + l2[2] * l2[2] + l2[3] * l2[3]);
        std::vector<double> dt(static_cast<std::size_t>(nc_)), alpha(static_cast<std::size_t>(nc_));
        double lo = 0.0, hi = 0.0;
        if (nc_ > 0 && cfdg_get_dt(ctx_, dt.data(), alpha.data()) == CFDG_OK) {
            lo = *std::min_element(dt.begin(), dt.end());
            hi = *std::max_element(dt.begin(), dt.end());
        }
        logf("iter %lld/%lld  wall %7.2fs  L2[mass]=%.4e  L2[mom]=%.4e  L2[energy]=%.4e  rel=%.3e  dt=[%.3e,%.3e]", it,
             static_cast<long long>(cfg_.max_iterations), wall, l2[0], mom, l2[4], rel, lo, hi);
    }

    // log_boundary_integrals (solver.hpp:826-848): the fluxes of the current state
    int boundary_integrals() {
        const auto np = static_cast<int32_t>(mesh_->m.patch_names.size());
        std::vector<double> mass(static_cast<std::size_t>(np)), energy(static_cast<std::size_t>(np));
        if (cfdg_eval_residual(ctx_) != CFDG_OK) return CFDG_ERR_STATE;
        const int rc = cfdg_boundary_integrals(ctx_, mass.data(), energy.data(), nullptr, np);
        if (rc != CFDG_OK) return rc;
        for (int32_t p = 0; p < np; ++p)
            logf("  patch %lld '%s': mass flux %+.6e kg/s, energy flux %+.6e W", static_cast<long long>(p),
                 mesh_->m.patch_names[static_cast<std::size_t>(p)].c_str(), mass[static_cast<std::size_t>(p)],
                 energy[static_cast<std::size_t>(p)]);
        return CFDG_OK;
    }

    static void linear_summary(const cfdg_iter_result& r) {
        const char* s = r.status == CFDG_CONVERGED ? "converged" : r.status == CFDG_DIVERGED ? "diverged" : "max-iterations";
        const double scale = r.initial_residual > 0.0 ? r.initial_residual : 1.0;  // ||b||: the step starts from zero
        std::printf("  [%s] %-14s iters %4d  |r|/|b| = %.3e\n", "BiCGSTAB", s, r.iterations, r.final_residual / scale);
        std::fflush(stdout);
    }
};

}  // namespace

extern "C" int cfdg_run_solver(const cfdg_solver_config* cfg, const cfdg_bc* bcs, int32_t n_bcs, const cfdg_mesh* mesh,
                               int32_t verbose, char* err, int32_t errlen) {
    if (!cfg || !bcs || !mesh) return CFDG_ERR_ARG;
    try {
        Run run(*cfg, mesh, verbose);
        const int rc = run.start(bcs, n_bcs, err, errlen);
        if (rc != CFDG_OK) return rc;
        return cfg->time_mode == CFDG_CFG_MODE_UNSTEADY ? run.dual_time(err, errlen) : run.steady(err, errlen);
    } catch (const std::exception& e) {
        put_msg(err, errlen, e.what());
        return CFDG_ERR_ARG;
    }
}
