// Driver that runs the REFERENCE's own configuration parsers
// (parse_solver_config, src/solver/compressible/config.cpp:28-288;
// parse_boundary_config, src/bc/config.cpp:141-289) on a file and prints the
// parsed structure, one "key value" pair per line (doubles with 17
// significant digits).  A rejected file ends in the reference's mpi::fatal,
// which prints its message on stderr and aborts.  Built by
// oracle/build_ref.sh into oracle/_ref/ (never committed); test
// infrastructure only.
//
// usage: ref_config solver <solver.toml>
//        ref_config bc <bc.toml> <patch name> ...
#include <mpi.h>

#include <cstdio>
#include <string>

#include "cfd/bc/config.hpp"
#include "cfd/mesh/localmesh.hpp"
#include "cfd/mpi/log.hpp"
#include "cfd/solver/compressible/config.hpp"

namespace {

void pd(const char* k, double v) { std::printf("%s %.17g\n", k, v); }
void pi(const char* k, long long v) { std::printf("%s %lld\n", k, v); }

}  // namespace

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    cfd::mpi::log_init(1);  // log_stat warnings go to stderr
    if (argc < 3) {
        std::fprintf(stderr, "usage: ref_config solver <file> | bc <file> <patch names...>\n");
        MPI_Finalize();
        return 2;
    }
    const std::string kind = argv[1], path = argv[2];
    if (kind == "solver") {
        const auto c = cfd::solver::compressible::parse_solver_config(path, MPI_COMM_WORLD);
        pi("eos", static_cast<long long>(c.flow.type));
        pd("gamma", c.flow.gamma);
        pd("gas_constant", c.flow.gas_constant);
        pd("prandtl", c.prandtl);
        pi("flow_model", static_cast<long long>(c.flow_model));
        pd("init_rho", c.init_rho);
        pd("init_p", c.init_p);
        for (int k = 0; k < 3; ++k) pd(("init_velocity" + std::to_string(k)).c_str(), c.init_velocity[k]);
        pi("gradient", static_cast<long long>(c.gradient));
        pi("flux", static_cast<long long>(c.flux));
        pi("reconstruction", static_cast<long long>(c.reconstruction));
        pi("limiter", static_cast<long long>(c.limiter));
        pd("venkat_k", c.limiter_venkat_k);
        pi("turbulence", c.turbulence.enabled ? 1 : 0);
        pd("nu_inf_ratio", c.turbulence.nu_inf_ratio);
        pi("max_distance_sweeps", c.turbulence.max_distance_sweeps);
        pd("distance_tolerance", c.turbulence.distance_tolerance);
        pi("scheme", static_cast<long long>(c.scheme));
        pd("cfl", c.cfl);
        pi("max_iterations", c.max_iterations);
        pd("residual_tolerance", c.residual_tolerance);
        pi("time_mode", static_cast<long long>(c.time_mode));
        pd("dt", c.dt);
        pi("max_time_steps", c.max_time_steps);
        pi("bdf_order", c.bdf_order);
        pi("solver", static_cast<long long>(c.linear_solver_params.type));
        pi("preconditioner", static_cast<long long>(c.linear_solver_params.precond_type));
        pd("rel_tol", c.linear_solver_params.relative_tolerance);
        pd("abs_tol", c.linear_solver_params.absolute_tolerance);
        pi("max_iter", c.linear_solver_params.max_iterations);
        pi("verbosity", static_cast<long long>(c.linear_solver_params.verbosity));
        std::printf("output_dir %s\n", c.output_dir.c_str());
        pi("field_interval", c.field_interval);
        pi("residual_interval", c.residual_interval);
    } else {
        cfd::mesh::MeshPart mp;
        for (int a = 3; a < argc; ++a) mp.patches.push_back({argv[a], "UserDefined"});
        const auto b = cfd::bc::parse_boundary_config(path, mp, MPI_COMM_WORLD);
        for (std::size_t p = 0; p < b.patches.size(); ++p) {
            const auto& d = b.patches[p];
            const std::string pre = "bc" + std::to_string(p) + ".";
            pi((pre + "type").c_str(), static_cast<long long>(d.type));
            pi((pre + "inflow_mode").c_str(), static_cast<long long>(d.inflow_mode));
            pd((pre + "p").c_str(), d.p);
            pd((pre + "t").c_str(), d.t);
            pd((pre + "tmp_grad").c_str(), d.tmp_grad);
            for (int k = 0; k < 3; ++k) pd((pre + "velocity" + std::to_string(k)).c_str(), d.velocity[k]);
            pd((pre + "mach").c_str(), d.mach);
            pd((pre + "alpha_deg").c_str(), d.alpha_deg);
            pd((pre + "beta_deg").c_str(), d.beta_deg);
            for (int k = 0; k < 3; ++k) pd((pre + "direction" + std::to_string(k)).c_str(), d.direction[k]);
        }
    }
    std::fflush(stdout);
    MPI_Finalize();
    return 0;
}
