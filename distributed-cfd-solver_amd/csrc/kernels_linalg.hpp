// Launch wrappers of the BSR / Krylov kernels (kernels_linalg.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace cfdg {

constexpr int kReduceBlocks = 1024;  // partial sums per reduction (fixed -> deterministic)

struct DotPairs {
    const double* a[5];
    const double* b[5];
};

void launch_lu_factor(int n, const int* diag, const double* values, double* lu, std::uint8_t* piv, int* singular,
                      hipStream_t s);
void launch_spmv(const BsrArgs& b, const double* values, const double* x, double* y, hipStream_t s);
void launch_sgs_group(const int* rows, int count, const BsrArgs& b, const double* values, const double* lu,
                      const std::uint8_t* piv, const double* r, double* z, hipStream_t s);
void launch_update_p(long long n, double beta, double nbo, const double* r, const double* v, double* p,
                     hipStream_t s);
void launch_axpy_copy(long long n, const double* x, double a, const double* z, double* y, hipStream_t s);
void launch_update_xr(long long n, double alpha, double omega, const double* ph, const double* sh, const double* sv,
                      const double* t, double* x, double* r, hipStream_t s);
int reduction_blocks(long long n);
// out[0..ndot) = dot products; partial needs kReduceBlocks * 5 doubles
void launch_dots(long long n, int ndot, const DotPairs& pr, double* partial, double* out, hipStream_t s);
void launch_res_sq(int nc, int ld, const double* res, double* partial, double* out, hipStream_t s);
void launch_rhs(int nc, int ld, const double* res, double* b, hipStream_t s);
void launch_state_update(int nc, int ld, const double* du, double* U, hipStream_t s);

}  // namespace cfdg
