// cfdg_solver: the reference's solver application on one GPU
// (apps/solver.cpp:20-105): the same command line, container and
// configuration files, run by the device path.
//
// Usage: cfdg_solver <mesh.h5> <solver.toml> <bc.toml> [-v|-vv]
//
// A fatal error prints the reference's mpi::fatal line (src/mpi/log.cpp:57-65,
// its wording included) on stderr and exits with status 1.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "cfdg.h"
#include "cfdg_front.h"

namespace {

[[noreturn]] void fatal(const std::string& what) {
    std::fprintf(stderr, "[Rank 0] CGNS Reader Fatal Error: %s\n", what.c_str());
    std::fflush(stderr);
    std::exit(EXIT_FAILURE);
}

int usage() {
    std::fprintf(stderr, "usage: cfdg_solver <mesh.h5> <solver.toml> <bc.toml> [-v|-vv]\n");
    return EXIT_FAILURE;
}

}  // namespace

int main(int argc, char** argv) {
    std::string mesh_file, solver_cfg, bc_cfg;
    int verbose = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-v" || a == "--verbose")
            verbose = 1;
        else if (a == "-vv")
            verbose = 2;
        else if (mesh_file.empty())
            mesh_file = a;
        else if (solver_cfg.empty())
            solver_cfg = a;
        else if (bc_cfg.empty())
            bc_cfg = a;
        else
            return usage();
    }
    if (mesh_file.empty() || solver_cfg.empty() || bc_cfg.empty()) return usage();

    std::fprintf(stderr, "solver: mesh=%s config=%s bc=%s ranks=%d\n", mesh_file.c_str(), solver_cfg.c_str(),
                 bc_cfg.c_str(), 1);
    const auto t0 = std::chrono::steady_clock::now();
    char err[4096] = {0};

    cfdg_mesh* mesh = nullptr;
    if (cfdg_container_read(mesh_file.c_str(), &mesh, err, sizeof err) != CFDG_OK) fatal(err);
    cfdg_mesh_view mv{};
    cfdg_mesh_get_view(mesh, &mv);
    if (verbose >= 1) {
        std::fprintf(stderr, "INFO: Mesh successfully loaded from Parallel HDF5: %s\n", mesh_file.c_str());
        std::fprintf(stderr, "mesh: %d cells, %d faces (%d interior), %d nodes, %d patches\n", mv.n_cells, mv.n_faces,
                     mv.n_inner_faces, mv.n_nodes, mv.n_patches);
    }

    cfdg_solver_config cfg{};
    if (cfdg_config_read_solver(solver_cfg.c_str(), &cfg, err, sizeof err) != CFDG_OK) fatal(err);
    std::vector<cfdg_bc> bcs(static_cast<std::size_t>(mv.n_patches));
    std::vector<char> warnings(1 << 16, '\0');
    if (cfdg_config_read_boundary(bc_cfg.c_str(), mesh, bcs.data(), mv.n_patches, warnings.data(),
                                  static_cast<int32_t>(warnings.size()), err, sizeof err) != CFDG_OK)
        fatal(err);
    if (verbose >= 1 && warnings[0] != '\0') std::fputs(warnings.data(), stderr);

    const int status = cfdg_run_solver(&cfg, bcs.data(), mv.n_patches, mesh, verbose, err, sizeof err);
    cfdg_mesh_free(mesh);
    if (status < 0) fatal(err);
    std::fprintf(stderr, "Total executional time = %.5f sec\n",
                 std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    return status == 0 ? EXIT_SUCCESS : EXIT_FAILURE;
}
