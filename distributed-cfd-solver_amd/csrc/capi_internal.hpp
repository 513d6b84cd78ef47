// Definitions shared by the C-ABI translation units (not part of the ABI).
#pragma once

#include <stdexcept>
#include <string>

#include "mesh.hpp"

struct cfdg_mesh {
    cfdg::Mesh m;
};

namespace cfdg {

// Error carrying a CFDG_ERR_* status code across the C boundary.
struct CfdgError : std::runtime_error {
    int code;
    CfdgError(int c, const std::string& w) : std::runtime_error(w), code(c) {}
};

}  // namespace cfdg
