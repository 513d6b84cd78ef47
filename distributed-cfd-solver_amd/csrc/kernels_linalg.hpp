// Launch wrappers of the BSR / Krylov kernels (kernels_linalg.hip).
//
// Matrix values are passed untyped with their storage kind (kStoreF64,
// kStoreF32, kStoreBF16 of kernels.hpp); vectors are typed (double or float)
// and fix the arithmetic type of the kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace cfdg {

constexpr int kReduceBlocks = 1024;  // partial sums per reduction (fixed -> deterministic)

// State block of the device-driven BiCGSTAB (doubles).  kKsStatus is 0 while
// the solve runs, then one of the codes below; every gated kernel returns at
// once when it is not 0.
enum : int {
    kKsStatus = 0, kKsIt, kKsMaxIt, kKsTol, kKsRho, kKsRhoOld, kKsAlpha, kKsOmega, kKsBeta, kKsNbo, kKsFirst,
    kKsOmegaNew, kKsRnNew, kKsRhoNew, kKsRnorm, kKsRnorm0, kKsReject, kKsCount = 24
};
constexpr double kKsConverged = 1.0, kKsDiverged = 2.0, kKsMaxIter = 3.0;

template <class T>
struct DotPairsT {
    const T* a[5];
    const T* b[5];
};
using DotPairs = DotPairsT<double>;

// Diagonal-block LU with pivoting, factorised in FP64 from the stored blocks
// and kept in FP64 (lu_f32 = 0, FP64 blocks only) or FP32 (lu_f32 = 1).
void launch_lu_factor(int n, const int* diag, const void* values, int storage, void* lu, int lu_f32,
                      std::uint8_t* piv, int* singular, hipStream_t s);
// gate: optional BiCGSTAB state block; the kernel does nothing once its status is set
template <class TV>
void launch_spmv(const BsrArgs& b, const void* values, int storage, const TV* x, TV* y, hipStream_t s,
                 const double* gate = nullptr);
// TV = double: FP64 blocks and FP64 LU.  TV = float: any storage, FP32 LU.
// lower (one flag per block, may be null = every block): only the blocks
// flagged 1 -- columns in an earlier group of the sweep -- are read.
template <class TV>
void launch_sgs_group(const int* rows, int count, const BsrArgs& b, const void* values, int storage, const void* lu,
                      const std::uint8_t* piv, const std::uint8_t* lower, const TV* r, TV* z, hipStream_t s,
                      const double* gate = nullptr);
template <class T>
void launch_update_p(long long n, double beta, double nbo, const T* r, const T* v, T* p, hipStream_t s);
template <class T>
void launch_axpy_copy(long long n, const T* x, double a, const T* z, T* y, hipStream_t s);
template <class T>
void launch_update_xr(long long n, double alpha, double omega, const T* ph, const T* sh, const T* sv, const T* t, T* x,
                      T* r, hipStream_t s);
void launch_cast_f64_to_f32(long long n, const double* src, float* dst, hipStream_t s);
void launch_cast_f32_to_f64(long long n, const float* src, double* dst, hipStream_t s);
void launch_add_f32(long long n, const float* e, double* x, hipStream_t s);
int reduction_blocks(long long n);
// out[0..ndot) = dot products accumulated in FP64; partial needs kReduceBlocks * 5 doubles
template <class T>
void launch_dots(long long n, int ndot, const DotPairsT<T>& pr, double* partial, double* out, hipStream_t s,
                 const double* gate = nullptr);
// device-driven BiCGSTAB: scalar steps on the state block ks and gated vector
// updates.  The start of each iteration (iteration limit, beta) is taken at
// the end of the step before it, so an iteration begins with its p update.
// Start of a solve from dots = (b, b) and, unless x_zero, (r, r) with r = b - A x:
// tolerance max(abs_tol, rel_tol ||b||), status, rho; with fallback, a guess
// whose residual exceeds ||b|| is marked rejected (launch_bicg_reject then
// restarts from zero).
void launch_bicg_init(double* ks, const double* dots, int x_zero, int fallback, double abs_tol, double rel_tol,
                      int max_iter, hipStream_t s);
template <class T>
void launch_bicg_reject(long long n, const double* ks, const T* b, T* x, T* r, hipStream_t s);
// dot products as launch_dots, gated by ks, then the scalar step that uses them
// in the reducing kernel: kStepAlpha (dots = (r0, v)), kStepOmega (dots =
// (t,t), (t,s), (r0,t), (r0,s), (s,s)) or kStepPost (dots = (r, r), the true
// residual norm, and the end of the iteration)
enum { kStepAlpha = 1, kStepOmega = 2, kStepPost = 3 };
template <class T>
void launch_dots_step(long long n, int ndot, const DotPairsT<T>& pr, double* partial, double* out, double* ks,
                      int step, hipStream_t s);
// the end of an iteration with the recurrence's residual norm
void launch_bicg_post(double* ks, hipStream_t s);
template <class T>
void launch_update_p_g(long long n, const double* ks, const T* r, const T* v, T* p, hipStream_t s);
template <class T>
void launch_axpy_alpha_g(long long n, const double* ks, const T* r, const T* v, T* sv, hipStream_t s);
// y = p (op 0, as launch_update_p_g) or s (op 1, as launch_axpy_alpha_g) over
// nrows block rows, and z_i = D_i^-1 y_i for the rows flagged in first (the SGS
// forward sweep's first group, factors lu / piv as launch_sgs_group takes them)
// The end of an iteration (x, r, checks) and, unless the solve ends there,
// the start of the next with p and the SGS forward sweep's first group of
// ph = M^-1 p, in one launch; the state is read from ks_in and written to
// ks_out (a second state block)
template <class TV>
void launch_xr_next(int nrows, const double* ks_in, double* ks_out, TV* ph, const TV* sh, const TV* sv, const TV* t,
                    TV* x, TV* r, const TV* v, TV* p, const std::uint8_t* first, int storage, const void* lu,
                    const std::uint8_t* piv, hipStream_t s);
template <class TV>
void launch_vec_sgs_first(int op, int nrows, const double* ks, const TV* r, const TV* v, TV* y,
                          const std::uint8_t* first, int storage, const void* lu, const std::uint8_t* piv, TV* z,
                          hipStream_t s);
template <class T>
void launch_update_xr_g(long long n, const double* ks, const T* ph, const T* sh, const T* sv, const T* t, T* x, T* r,
                        hipStream_t s);
void launch_res_sq(int nc, int ld, const double* res, double* partial, double* out, hipStream_t s);
// l2 (5 doubles, device) = residual norms of res (5 planes of ld, nc cells,
// device), with the partial sums in the library's static scratch array: no
// device allocation, no host synchronisation; stream-ordered on s.
hipError_t residual_norms_noalloc(int nc, int ld, const double* res, double* l2, hipStream_t s);
void launch_rhs(int nc, int ld, const double* res, double* b, hipStream_t s);
void launch_state_update(int nc, int ld, const double* du, double* U, hipStream_t s);

}  // namespace cfdg
