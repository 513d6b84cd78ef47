// Definitions shared by the C-ABI translation units (not part of the ABI).
#pragma once

#include <stdexcept>
#include <string>

#include <memory>

#include "mesh.hpp"
#include "setup.hpp"

struct cfdg_mesh {
    cfdg::Mesh m;
    // cell connectivity built by cfdg_mesh_connectivity (dropped when the
    // cells are renumbered); contexts created from the mesh use it
    std::shared_ptr<const cfdg::MeshConnectivity> conn;
};

namespace cfdg {

// Error carrying a CFDG_ERR_* status code across the C boundary.
struct CfdgError : std::runtime_error {
    int code;
    CfdgError(int c, const std::string& w) : std::runtime_error(w), code(c) {}
};

}  // namespace cfdg
