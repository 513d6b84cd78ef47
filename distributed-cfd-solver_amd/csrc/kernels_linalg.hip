// Block-sparse (5x5 BSR) linear-algebra kernels for the implicit step (gfx950).
//
//   k_lu_factor   diagonal-block LU with partial pivoting (bsr_helpers.hpp:278-321)
//   k_spmv        y = A x, five lanes per block row (bsr_helpers.hpp:139-203)
//   k_sgs_group   one Gauss-Seidel row update for every row of one row group
//                 (bsr_helpers.hpp:390-435 for a single row), eight lanes per
//                 row; a sweep launches the groups in order (colours, or
//                 dependency levels)
//   vector ops    the BiCGSTAB updates of bicgstab.cpp:97-177, fused
//   reductions    dot products: per-block partial sums in a fixed order, then
//                 one block adds the partials in a fixed order -> the result
//                 is identical from run to run (no atomics)
//
// Precision forms: the kernels are templated on the stored block type TS
// (double, float, bf16), the diagonal-LU type TL (double, float) and the
// vector / arithmetic type TV (double, float).  Every combination that uses
// double for all three performs the FP64 arithmetic of the reference in its
// order; dot products always accumulate in double.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <stdexcept>

#include "kernels.hpp"
#include "kernels_linalg.hpp"

// Scratch of the allocation-free residual norms (residual_norms_noalloc):
// kReduceBlocks partial 5-vectors, 40 KiB per GPU, held until the process
// ends.  Host code never passes this name to a kernel: it is a host-side
// stand-in.  The device address comes from hipGetSymbolAddress, once per
// device, with that device current (noalloc_scratch below).
static __device__ __align__(256) double scratch[cfdg::kReduceBlocks * 5];

namespace cfdg {

namespace {

constexpr int kBlock = 256;

inline dim3 grid_for(long long n) { return dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)); }

[[noreturn]] void throw_bad_combination(const char* what) { throw std::invalid_argument(what); }

template <class TS, class TL>
__global__ void __launch_bounds__(kBlock) k_lu_factor(int n, const int* diag, const TS* values, TL* lu,
                                                      std::uint8_t* piv, int* singular) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // fully unrolled so that the 25 entries stay in registers: the row swap is
    // a select over the (compile-time) candidate rows
    double a[25];
    const TS* src = values + static_cast<long long>(diag[i]) * BlockLayout<TS>::block;
#pragma unroll
    for (int k = 0; k < 25; ++k) a[k] = to_f64(src[(k / 5) * BlockLayout<TS>::row + k % 5]);
    int p[5] = {0, 1, 2, 3, 4};
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        int mrow = k;
        double mx = fabs(a[k * 5 + k]);
#pragma unroll
        for (int r = k + 1; r < 5; ++r) {
            const double v = fabs(a[r * 5 + k]);
            if (v > mx) {
                mx = v;
                mrow = r;
            }
        }
        p[k] = mrow;
        ok = ok && (mx != 0.0);
#pragma unroll
        for (int r = k + 1; r < 5; ++r) {
            const bool sw = (mrow == r);
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                const double tk = a[k * 5 + j], tr = a[r * 5 + j];
                a[k * 5 + j] = sw ? tr : tk;
                a[r * 5 + j] = sw ? tk : tr;
            }
        }
        const double inv_d = 1.0 / a[k * 5 + k];
#pragma unroll
        for (int r = k + 1; r < 5; ++r) {
            const double l = a[r * 5 + k] * inv_d;
            a[r * 5 + k] = l;
#pragma unroll
            for (int j = k + 1; j < 5; ++j) a[r * 5 + j] -= l * a[k * 5 + j];
        }
    }
    if (!ok) {
        atomicMax(singular, i + 1);  // integer flag: largest singular row + 1
        return;
    }
    TL* dst = lu + static_cast<long long>(i) * 25;
#pragma unroll
    for (int k = 0; k < 25; ++k) dst[k] = cvt<TL>(a[k]);
#pragma unroll
    for (int k = 0; k < 5; ++k) piv[static_cast<long long>(i) * 5 + k] = static_cast<std::uint8_t>(p[k]);
}

// Row ii of stored block k in the arithmetic type.  bfloat16 rows are read
// as three 4-byte words (the row is 4-byte aligned), one wide load.
template <class TV, class TS>
__device__ __forceinline__ void block_row(const TS* values, long long k, int ii, TV a[5]) {
    const TS* p = values + k * BlockLayout<TS>::block + ii * BlockLayout<TS>::row;
#pragma unroll
    for (int q = 0; q < 5; ++q) a[q] = load_as<TV>(p[q]);
}
template <class TV>
__device__ __forceinline__ void block_row(const bf16* values, long long k, int ii, TV a[5]) {
    const std::uint32_t* p =
        reinterpret_cast<const std::uint32_t*>(values + k * BlockLayout<bf16>::block + ii * BlockLayout<bf16>::row);
    const std::uint32_t w0 = p[0], w1 = p[1], w2 = p[2];
    a[0] = static_cast<TV>(f32_from_bits(w0 << 16));
    a[1] = static_cast<TV>(f32_from_bits(w0 & 0xffff0000u));
    a[2] = static_cast<TV>(f32_from_bits(w1 << 16));
    a[3] = static_cast<TV>(f32_from_bits(w1 & 0xffff0000u));
    a[4] = static_cast<TV>(f32_from_bits(w2 << 16));
}

// y = A x with five lanes per block row: lane ii forms
// acc_ii = sum_k (sum_jj a_k[ii][jj] x_col(k)[jj]) in column order.
template <class TS, class TV>
__global__ void __launch_bounds__(kBlock) k_spmv(int n, const int* row_ptr, const int* cols, const TS* values,
                                                 const TV* x, TV* y, const double* gate) {
    if (gate && gate[kKsStatus] != 0.0) return;
    const long long t = static_cast<long long>(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
    const long long row = t / 5;
    const int ii = static_cast<int>(t - row * 5);
    if (row >= n) return;
    TV acc = 0;
    const int kb = row_ptr[row], ke = row_ptr[row + 1];
    for (int k = kb; k < ke; ++k) {
        TV a[5];
        block_row(values, k, ii, a);
        const TV* xb = x + static_cast<long long>(cols[k]) * 5;
        TV s = 0;
        s += a[0] * xb[0];
        s += a[1] * xb[1];
        s += a[2] * xb[2];
        s += a[3] * xb[3];
        s += a[4] * xb[4];
        acc += s;
    }
    y[row * 5 + ii] = acc;
}

// One Gauss-Seidel update for each row of a group: z_i = LU_i^-1 (r_i - sum_{k != d} A_ik z_k).
// Five lanes per row, twelve rows per wave (lanes 60-63 idle): lane ii forms
// component ii of the sum over the row's off-diagonal blocks, reading block
// row ii (5 contiguous values), so the five lanes of a row read each 5x5 block
// as one contiguous run; the five sums are then exchanged within the row's
// five lanes and every lane performs the same 5x5 LU solve
// (lu_solve_block<5>, bsr_helpers.hpp:325-353) and keeps its component.  The
// arithmetic and its order per component are those of one thread per row.
constexpr int kSgsRowsPerWave = 12;

// Component ii of the solution of (LU) x = s with partial pivoting
// (lu_solve_block<5>, bsr_helpers.hpp:325-353): row swaps, forward then
// backward substitution over the five values every lane of a row holds.
template <class TV>
__device__ __forceinline__ TV lu_solve_component(TV s[5], const TV Lr[25], const int pv[5], int ii) {
    for (int k = 0; k < 5; ++k) {
        const int pk = pv[k];
        if (pk != k) {
            TV tmp = s[k];
            TV sp = s[0];
            for (int q = 1; q < 5; ++q) sp = (q == pk) ? s[q] : sp;
            s[k] = sp;
            for (int q = 0; q < 5; ++q) s[q] = (q == pk) ? tmp : s[q];
        }
    }
    for (int a = 1; a < 5; ++a) {
        TV acc2 = s[a];
        for (int b = 0; b < a; ++b) acc2 -= Lr[a * 5 + b] * s[b];
        s[a] = acc2;
    }
    for (int a = 4; a >= 0; --a) {
        TV acc2 = s[a];
        for (int b = a + 1; b < 5; ++b) acc2 -= Lr[a * 5 + b] * s[b];
        s[a] = acc2 / Lr[a * 5 + a];
    }
    TV out = s[0];
    for (int q = 1; q < 5; ++q) out = (q == ii) ? s[q] : out;
    return out;
}

template <class TS, class TL, class TV>
__global__ void __launch_bounds__(kBlock) k_sgs_group(const int* rows, int count, const int* row_ptr, const int* cols,
                                                      const int* diag, const TS* values, const TL* lu,
                                                      const std::uint8_t* piv, const std::uint8_t* lower, const TV* r,
                                                      TV* z, const double* gate) {
    if (gate && gate[kKsStatus] != 0.0) return;
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int rw = lane / 5;  // row of the wave, 12 = idle lanes
    const int ii = lane - rw * 5;
    const long long wave = (static_cast<long long>(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x) >> 6;
    const long long slot = wave * kSgsRowsPerWave + rw;
    const bool has_row = rw < kSgsRowsPerWave && slot < count;  // uniform over the row's five lanes
    const int i = has_row ? rows[slot] : 0;
    TV acc = 0;
    // FP32 factors: the row's LU factors and pivots are requested before the
    // off-diagonal sweep, so their latency overlaps it instead of following the
    // exchange.  FP64 factors are loaded after it (50 more live registers cost
    // the FP64 kernel more occupancy than the overlap gains).
    constexpr bool kEarlyLu = sizeof(TL) == sizeof(float);
    TV Lr[25];
    int pv[5];
    const TL* L = lu + static_cast<long long>(i) * 25;
    const std::uint8_t* p = piv + static_cast<long long>(i) * 5;
    if (has_row && ii < 5) {
        if constexpr (kEarlyLu) {
            for (int q = 0; q < 25; ++q) Lr[q] = load_as<TV>(L[q]);
            for (int q = 0; q < 5; ++q) pv[q] = p[q];
        }
        acc = r[static_cast<long long>(i) * 5 + ii];
        const int d = diag[i];
        const int kb = row_ptr[i], ke = row_ptr[i + 1];
        for (int k = kb; k < ke; ++k) {
            // lower: the sweep reads only the blocks whose column lies in an
            // earlier group (the others multiply values not computed yet)
            if (k == d || (lower && !lower[k])) continue;
            TV a[5];
            block_row(values, k, ii, a);
            const TV* zb = z + static_cast<long long>(cols[k]) * 5;
            TV tt = 0;
            tt += a[0] * zb[0];
            tt += a[1] * zb[1];
            tt += a[2] * zb[2];
            tt += a[3] * zb[3];
            tt += a[4] * zb[4];
            acc -= tt;
        }
    }
    const int base = rw * 5;  // idle lanes (rw = 12) read lanes 60-63 and discard
    TV s[5];
    for (int q = 0; q < 5; ++q) s[q] = __shfl(acc, (base + q) & 63, 64);
    if (!has_row || ii >= 5) return;
    if constexpr (!kEarlyLu) {
        for (int q = 0; q < 25; ++q) Lr[q] = load_as<TV>(L[q]);
        for (int q = 0; q < 5; ++q) pv[q] = p[q];
    }
    z[static_cast<long long>(i) * 5 + ii] = lu_solve_component(s, Lr, pv, ii);
}

// p = ((p * beta) + (-beta * omega) v) + r           (bicgstab.cpp:104-109)
template <class T>
__global__ void __launch_bounds__(kBlock) k_update_p(long long n, double beta, double nbo, const T* r, const T* v,
                                                     T* p) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T t = p[i] * static_cast<T>(beta);
    t += static_cast<T>(nbo) * v[i];
    t += static_cast<T>(1.0) * r[i];
    p[i] = t;
}

// y = x + a z
template <class T>
__global__ void __launch_bounds__(kBlock) k_axpy_copy(long long n, const T* x, double a, const T* z, T* y) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = x[i] + static_cast<T>(a) * z[i];
}

// x = (x + alpha ph) + omega sh;  r = s + (-omega) t   (bicgstab.cpp:145-149)
template <class T>
__global__ void __launch_bounds__(kBlock) k_update_xr(long long n, double alpha, double omega, const T* ph,
                                                      const T* sh, const T* s, const T* t, T* x, T* r) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T xv = x[i];
    xv += static_cast<T>(alpha) * ph[i];
    xv += static_cast<T>(omega) * sh[i];
    x[i] = xv;
    T rv = s[i];
    rv += static_cast<T>(-omega) * t[i];
    r[i] = rv;
}

// --- device-driven BiCGSTAB --------------------------------------------------
// The scalars of the recurrence live in the state block ks (kernels_linalg.hpp)
// and one-thread kernels update them from the dot products, so a batch of
// iterations is enqueued without host round trips.  Every kernel of an
// iteration returns at once when ks[kKsStatus] is no longer "running", so the
// iterations enqueued after convergence or breakdown change nothing and the
// result is the one the per-iteration host loop of bicgstab.cpp:42-201 gives.

// Start of an iteration (bicgstab.cpp:97-109): iteration limit, it += 1,
// beta = (rho / rho_old)(alpha / omega) from the second iteration on.  Run
// at the end of the step before it (the start of the solve, or the end of the
// previous iteration), since nothing in between reads or writes the state.
__device__ void bicg_begin(double* ks) {
    if (ks[kKsIt] >= ks[kKsMaxIt]) {
        ks[kKsStatus] = kKsMaxIter;
        return;
    }
    ks[kKsIt] += 1.0;
    ks[kKsFirst] = (ks[kKsIt] == 1.0) ? 1.0 : 0.0;
    if (ks[kKsFirst] == 0.0) {
        const double beta = (ks[kKsRho] / ks[kKsRhoOld]) * (ks[kKsAlpha] / ks[kKsOmega]);
        if (!isfinite(beta)) {
            ks[kKsStatus] = kKsDiverged;
            return;
        }
        ks[kKsBeta] = beta;
        ks[kKsNbo] = -beta * ks[kKsOmega];
    }
}

// Start of a solve (bicgstab.cpp:56-92): norms, tolerance, initial status,
// then the start of the first iteration.
__global__ void k_bicg_init(double* ks, const double* dots, int x_zero, int fallback, double abs_tol,
                            double rel_tol, int max_iter) {
    const double bnorm = sqrt(dots[0]);
    double rnorm = x_zero ? bnorm : sqrt(dots[1]);
    double reject = 0.0;
    if (fallback && !(rnorm <= bnorm)) {
        reject = 1.0;
        rnorm = bnorm;
    }
    const double scale = (bnorm > 0.0) ? bnorm : 1.0;
    const double tol = dmax(abs_tol, rel_tol * scale);
    for (int k = 0; k < kKsCount; ++k) ks[k] = 0.0;
    ks[kKsMaxIt] = static_cast<double>(max_iter);
    ks[kKsTol] = tol;
    ks[kKsRho] = rnorm * rnorm;
    ks[kKsRhoOld] = 1.0;
    ks[kKsAlpha] = 1.0;
    ks[kKsOmega] = 1.0;
    ks[kKsRnorm] = rnorm;
    ks[kKsRnorm0] = rnorm;
    ks[kKsReject] = reject;
    if (!isfinite(rnorm))
        ks[kKsStatus] = kKsDiverged;
    else if (rnorm <= tol)
        ks[kKsStatus] = kKsConverged;
    else
        bicg_begin(ks);
}

// A rejected guess restarts from zero: x = 0, r = b.
template <class T>
__global__ void __launch_bounds__(kBlock) k_bicg_reject(long long n, const double* ks, const T* b, T* x, T* r) {
    if (ks[kKsReject] == 0.0) return;
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = static_cast<T>(0);
    r[i] = b[i];
}

// gamma = (r0, v); alpha = rho / gamma (bicgstab.cpp:113-120)
__device__ void bicg_alpha(double* ks, const double* dots) {
    const double gamma = dots[0];
    if (!isfinite(gamma) || gamma == 0.0) {
        ks[kKsStatus] = kKsDiverged;
        return;
    }
    ks[kKsAlpha] = ks[kKsRho] / gamma;
}

// omega = (t, s) / (t, t); the recurrence norm of the new r and rho_new
// (bicgstab.cpp:126-160); dots = (t,t), (t,s), (r0,t), (r0,s), (s,s)
__device__ void bicg_omega(double* ks, const double* dots) {
    const double tt = dots[0], ts = dots[1], r0t = dots[2], r0s = dots[3], ss = dots[4];
    if (!isfinite(ts) || !isfinite(tt) || tt == 0.0) {
        ks[kKsStatus] = kKsDiverged;
        return;
    }
    const double omega_new = ts / tt;
    const double rn2 = ss - 2.0 * omega_new * ts + omega_new * omega_new * tt;
    ks[kKsOmegaNew] = omega_new;
    ks[kKsRnNew] = sqrt(rn2 > 0.0 ? rn2 : 0.0);
    ks[kKsRhoNew] = r0s - omega_new * r0t;
}

// After x and r are updated (bicgstab.cpp:160-195): breakdown checks, the
// true residual norm when rr is given (every 32nd iteration), convergence;
// then the start of the next iteration.
__device__ void bicg_post(double* ks, const double* rr) {
    double rnorm = ks[kKsRnNew];
    const double omega_new = ks[kKsOmegaNew];
    ks[kKsRnorm] = rnorm;
    if (!isfinite(rnorm) || !isfinite(omega_new) || omega_new == 0.0) {
        ks[kKsStatus] = kKsDiverged;
        return;
    }
    if (rr) rnorm = sqrt(rr[0]);
    ks[kKsRnorm] = rnorm;
    if (rnorm <= ks[kKsTol]) {
        ks[kKsStatus] = kKsConverged;
        return;
    }
    const double rho_new = ks[kKsRhoNew];
    if (!isfinite(rho_new) || rho_new == 0.0) {
        ks[kKsStatus] = kKsDiverged;
        return;
    }
    ks[kKsRhoOld] = ks[kKsRho];
    ks[kKsRho] = rho_new;
    ks[kKsOmega] = omega_new;
    bicg_begin(ks);
}

__global__ void k_bicg_post(double* ks) {
    if (ks[kKsStatus] != 0.0) return;
    bicg_post(ks, nullptr);
}

// p = r on the first iteration, else p = ((p beta) + (-beta omega) v) + r
template <class T>
__device__ __forceinline__ T p_entry_of(double first, double beta, double nbo, T ri, const T* v, const T* p,
                                        long long i) {
    if (first != 0.0) return ri;
    T t = p[i] * static_cast<T>(beta);
    t += static_cast<T>(nbo) * v[i];
    t += static_cast<T>(1.0) * ri;
    return t;
}
template <class T>
__device__ __forceinline__ T p_entry(const double* ks, const T* r, const T* v, const T* p, long long i) {
    return p_entry_of(ks[kKsFirst], ks[kKsBeta], ks[kKsNbo], r[i], v, p, i);
}

// x = (x + alpha ph) + omega sh, written; returns r = s + (-omega) t
// (bicgstab.cpp:145-149)
template <class T>
__device__ __forceinline__ T xr_entry(const double* ks, const T* ph, const T* sh, const T* sv, const T* t, T* x,
                                      long long i) {
    const double alpha = ks[kKsAlpha], omega = ks[kKsOmegaNew];
    T xv = x[i];
    xv += static_cast<T>(alpha) * ph[i];
    xv += static_cast<T>(omega) * sh[i];
    x[i] = xv;
    T rv = sv[i];
    rv += static_cast<T>(-omega) * t[i];
    return rv;
}

// s = r + (-alpha) v
template <class T>
__device__ __forceinline__ T s_entry(const double* ks, const T* r, const T* v, long long i) {
    return r[i] + static_cast<T>(-ks[kKsAlpha]) * v[i];
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_update_p_g(long long n, const double* ks, const T* r, const T* v, T* p) {
    if (ks[kKsStatus] != 0.0) return;
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    p[i] = p_entry(ks, r, v, p, i);
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_axpy_alpha_g(long long n, const double* ks, const T* r, const T* v,
                                                         T* sv) {
    if (ks[kKsStatus] != 0.0) return;
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sv[i] = s_entry(ks, r, v, i);
}

// The SGS forward sweep's first group reads no off-diagonal block: its rows
// solve D_i z_i = y_i with the diagonal factors alone.  These kernels form a
// Krylov vector y -- p (kOp 0) or s (kOp 1), entry for entry as
// k_update_p_g / k_axpy_alpha_g -- and for the rows flagged in `first` also
// that solve, in k_sgs_group's lane layout and arithmetic, so z leaves them
// as the sweep's first group would and the sweep starts at its second group.
template <int kOp, class TL, class TV>
__global__ void __launch_bounds__(kBlock) k_vec_sgs_first(int nrows, const double* ks, const TV* r, const TV* v,
                                                          TV* y, const std::uint8_t* first, const TL* lu,
                                                          const std::uint8_t* piv, TV* z) {
    if (ks[kKsStatus] != 0.0) return;
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int rw = lane / 5;
    const int ii = lane - rw * 5;
    const long long wave = (static_cast<long long>(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x) >> 6;
    const long long row = wave * kSgsRowsPerWave + rw;
    const bool has_row = rw < kSgsRowsPerWave && row < nrows;  // uniform over the row's five lanes
    const long long e = row * 5 + ii;
    TV acc = 0;
    bool solve = false;
    if (has_row) {
        acc = (kOp == 0) ? p_entry(ks, r, v, y, e) : s_entry(ks, r, v, e);
        y[e] = acc;
        solve = first[row] != 0;
    }
    const int base = rw * 5;
    TV s[5];
    for (int q = 0; q < 5; ++q) s[q] = __shfl(acc, (base + q) & 63, 64);
    if (!solve) return;
    TV Lr[25];
    int pv[5];
    const TL* L = lu + row * 25;
    const std::uint8_t* p = piv + row * 5;
    for (int q = 0; q < 25; ++q) Lr[q] = load_as<TV>(L[q]);
    for (int q = 0; q < 5; ++q) pv[q] = p[q];
    z[e] = lu_solve_component(s, Lr, pv, ii);
}

// x = (x + alpha ph) + omega sh;  r = s + (-omega) t
// The end of one iteration and the start of the next in one launch: x and r
// as k_update_xr_g; the end-of-iteration checks and the next iteration's
// start (bicg_post without the true residual, then bicg_begin) formed by
// thread 0 of every block from the state ks_in alone (so every block takes
// the same decision), block 0 writing the result to ks_out; then, unless the
// solve has ended, p and the SGS forward sweep's first group of M^-1 p as
// k_vec_sgs_first.  ph is read (x update) and written (z = M^-1 p, first group)
// entry by entry by the same lane.
template <class TL, class TV>
__global__ void __launch_bounds__(kBlock) k_xr_next(int nrows, const double* ks_in, double* ks_out, TV* ph,
                                                    const TV* sh, const TV* sv, const TV* t, TV* x, TV* r,
                                                    const TV* v, TV* p, const std::uint8_t* first, const TL* lu,
                                                    const std::uint8_t* piv) {
    __shared__ double next[4];  // status, first, beta, nbo of the next iteration
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (ks_in[kKsStatus] != 0.0) {  // ended before: carry the state over
        if (lead)
            for (int k = 0; k < kKsCount; ++k) ks_out[k] = ks_in[k];
        return;
    }
    if (threadIdx.x == 0) {
        double st[kKsCount];
        for (int k = 0; k < kKsCount; ++k) st[k] = ks_in[k];
        bicg_post(st, nullptr);
        next[0] = st[kKsStatus];
        next[1] = st[kKsFirst];
        next[2] = st[kKsBeta];
        next[3] = st[kKsNbo];
        if (lead)
            for (int k = 0; k < kKsCount; ++k) ks_out[k] = st[k];
    }
    __syncthreads();
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int rw = lane / 5;
    const int ii = lane - rw * 5;
    const long long wave = (static_cast<long long>(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x) >> 6;
    const long long row = wave * kSgsRowsPerWave + rw;
    const bool has_row = rw < kSgsRowsPerWave && row < nrows;  // uniform over the row's five lanes
    const long long e = row * 5 + ii;
    const bool go = next[0] == 0.0;  // uniform over the grid
    TV acc = 0;
    bool solve = false;
    if (has_row) {
        const TV rv = xr_entry(ks_in, ph, sh, sv, t, x, e);
        r[e] = rv;
        if (go) {
            acc = p_entry_of(next[1], next[2], next[3], rv, v, p, e);
            p[e] = acc;
            solve = first[row] != 0;
        }
    }
    if (!go) return;
    const int base = rw * 5;
    TV s[5];
    for (int q = 0; q < 5; ++q) s[q] = __shfl(acc, (base + q) & 63, 64);
    if (!solve) return;
    TV Lr[25];
    int pv[5];
    const TL* L = lu + row * 25;
    const std::uint8_t* pp = piv + row * 5;
    for (int q = 0; q < 25; ++q) Lr[q] = load_as<TV>(L[q]);
    for (int q = 0; q < 5; ++q) pv[q] = pp[q];
    ph[e] = lu_solve_component(s, Lr, pv, ii);
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_update_xr_g(long long n, const double* ks, const T* ph, const T* sh,
                                                        const T* sv, const T* t, T* x, T* r) {
    if (ks[kKsStatus] != 0.0) return;
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    r[i] = xr_entry(ks, ph, sh, sv, t, x, i);
}

// dst = src converted (defect correction: FP64 residual -> FP32 inner right-hand side)
template <class TD, class TS>
__global__ void __launch_bounds__(kBlock) k_cast(long long n, const TS* src, TD* dst) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[i] = static_cast<TD>(src[i]);
}

// x += e with e in FP32 (defect correction update, in FP64)
__global__ void __launch_bounds__(kBlock) k_add_f32(long long n, const float* e, double* x) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = x[i] + static_cast<double>(e[i]);
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Block-level sum in a fixed tree order; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) s += lds[w];
    }
    __syncthreads();
    return s;
}

// Up to five dot products of flat vectors in one pass; partial[b * 5 + j].
template <class T>
__global__ void __launch_bounds__(kBlock) k_dots(long long n, int ndot, DotPairsT<T> pr, double* partial,
                                                 const double* gate) {
    if (gate && gate[kKsStatus] != 0.0) return;
    __shared__ double lds[kBlock / 64];
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        for (int j = 0; j < 5; ++j)
            if (j < ndot) acc[j] += static_cast<double>(pr.a[j][i]) * static_cast<double>(pr.b[j][i]);
    }
    for (int j = 0; j < 5; ++j) {
        if (j >= ndot) break;
        const double s = block_sum(acc[j], lds);
        if (threadIdx.x == 0) partial[blockIdx.x * 5 + j] = s;
    }
}

// Sum of squares of the five residual planes (solver.hpp:740-756)
__global__ void __launch_bounds__(kBlock) k_res_sq(int nc, int ld, const double* res, double* partial) {
    __shared__ double lds[kBlock / 64];
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += stride) {
        for (int v = 0; v < 5; ++v) {
            const double r = res[static_cast<long long>(v) * ld + i];
            acc[v] += r * r;
        }
    }
    for (int v = 0; v < 5; ++v) {
        const double s = block_sum(acc[v], lds);
        if (threadIdx.x == 0) partial[blockIdx.x * 5 + v] = s;
    }
}

// Adds nblocks partial 5-vectors: thread t sums blocks t, t + 256, ... in
// order, then the block tree adds the 256 thread sums in a fixed order.
__global__ void __launch_bounds__(kBlock) k_reduce_partials(int nblocks, int ncomp, const double* partial,
                                                            double* out, const double* gate) {
    if (gate && gate[kKsStatus] != 0.0) return;
    __shared__ double lds[kBlock / 64];
    for (int j = 0; j < ncomp; ++j) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += kBlock) s += partial[b * 5 + j];
        s = block_sum(s, lds);
        if (threadIdx.x == 0) out[j] = s;
    }
}

// k_reduce_partials followed, in thread 0, by the scalar step of the
// BiCGSTAB recurrence that consumes the sums (alpha, omega, or the end of an
// iteration with the true residual norm), so the step needs no launch of its
// own.  The block reads the state at its start, before thread 0 writes it.
template <int kStep>
__global__ void __launch_bounds__(kBlock) k_reduce_step(int nblocks, int ncomp, const double* partial, double* out,
                                                        double* ks) {
    if (ks[kKsStatus] != 0.0) return;
    __shared__ double lds[kBlock / 64];
    for (int j = 0; j < ncomp; ++j) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += kBlock) s += partial[b * 5 + j];
        s = block_sum(s, lds);
        if (threadIdx.x == 0) out[j] = s;
    }
    if (threadIdx.x != 0) return;
    if constexpr (kStep == kStepAlpha) bicg_alpha(ks, out);
    if constexpr (kStep == kStepOmega) bicg_omega(ks, out);
    if constexpr (kStep == kStepPost) bicg_post(ks, out);
}

// The partial sums of k_res_sq added in k_reduce_partials' order, then
// l2[v] = sqrt(sum_v * scale) (solver.hpp:740-756) written to device memory.
__global__ void __launch_bounds__(kBlock) k_norms_finish(int nblocks, const double* partial, double scale,
                                                         double* l2) {
    __shared__ double lds[kBlock / 64];
    for (int j = 0; j < 5; ++j) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += kBlock) s += partial[b * 5 + j];
        s = block_sum(s, lds);
        if (threadIdx.x == 0) l2[j] = sqrt(s * scale);
    }
}

// b = -R (AoS), the implicit right-hand side (implicit_euler.hpp:81-92)
__global__ void __launch_bounds__(kBlock) k_rhs(int nc, int ld, const double* res, double* b) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    for (int v = 0; v < 5; ++v) b[static_cast<long long>(c) * 5 + v] = -res[static_cast<long long>(v) * ld + c];
}

// U = U + du (implicit_euler.hpp:53-61)
__global__ void __launch_bounds__(kBlock) k_state_update(int nc, int ld, const double* du, double* U) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    for (int v = 0; v < 5; ++v) {
        double* u = U + static_cast<long long>(v) * ld + c;
        *u = *u + du[static_cast<long long>(c) * 5 + v];
    }
}

}  // namespace

void launch_lu_factor(int n, const int* diag, const void* values, int storage, void* lu, int lu_f32,
                      std::uint8_t* piv, int* singular, hipStream_t s) {
    if (n <= 0) return;
    const dim3 g = grid_for(n), b(kBlock);
    if (!lu_f32) {
        if (storage != kStoreF64) throw_bad_combination("FP64 LU needs FP64 blocks");
        hipLaunchKernelGGL((k_lu_factor<double, double>), g, b, 0, s, n, diag, static_cast<const double*>(values),
                           static_cast<double*>(lu), piv, singular);
    } else if (storage == kStoreF32) {
        hipLaunchKernelGGL((k_lu_factor<float, float>), g, b, 0, s, n, diag, static_cast<const float*>(values),
                           static_cast<float*>(lu), piv, singular);
    } else if (storage == kStoreBF16) {
        hipLaunchKernelGGL((k_lu_factor<bf16, float>), g, b, 0, s, n, diag, static_cast<const bf16*>(values),
                           static_cast<float*>(lu), piv, singular);
    } else {
        hipLaunchKernelGGL((k_lu_factor<double, float>), g, b, 0, s, n, diag, static_cast<const double*>(values),
                           static_cast<float*>(lu), piv, singular);
    }
}

template <class TV>
void launch_spmv(const BsrArgs& b, const void* values, int storage, const TV* x, TV* y, hipStream_t s,
                 const double* gate) {
    if (b.n <= 0) return;
    const dim3 g = grid_for(5LL * b.n), blk(kBlock);
    if (storage == kStoreF32)
        hipLaunchKernelGGL((k_spmv<float, TV>), g, blk, 0, s, b.n, b.row_ptr, b.cols,
                           static_cast<const float*>(values), x, y, gate);
    else if (storage == kStoreBF16)
        hipLaunchKernelGGL((k_spmv<bf16, TV>), g, blk, 0, s, b.n, b.row_ptr, b.cols, static_cast<const bf16*>(values),
                           x, y, gate);
    else
        hipLaunchKernelGGL((k_spmv<double, TV>), g, blk, 0, s, b.n, b.row_ptr, b.cols,
                           static_cast<const double*>(values), x, y, gate);
}
template void launch_spmv<double>(const BsrArgs&, const void*, int, const double*, double*, hipStream_t,
                                  const double*);
template void launch_spmv<float>(const BsrArgs&, const void*, int, const float*, float*, hipStream_t, const double*);

template <class TV>
void launch_sgs_group(const int* rows, int count, const BsrArgs& b, const void* values, int storage, const void* lu,
                      const std::uint8_t* piv, const std::uint8_t* lower, const TV* r, TV* z, hipStream_t s,
                      const double* gate) {
    if (count <= 0) return;
    const dim3 g = grid_for((static_cast<long long>(count) + kSgsRowsPerWave - 1) / kSgsRowsPerWave * 64), blk(kBlock);
    if constexpr (sizeof(TV) == sizeof(double)) {
        if (storage != kStoreF64) throw_bad_combination("FP64 SGS needs FP64 blocks");
        hipLaunchKernelGGL((k_sgs_group<double, double, double>), g, blk, 0, s, rows, count, b.row_ptr, b.cols, b.diag,
                           static_cast<const double*>(values), static_cast<const double*>(lu), piv, lower, r, z, gate);
    } else {
        if (storage == kStoreF32)
            hipLaunchKernelGGL((k_sgs_group<float, float, float>), g, blk, 0, s, rows, count, b.row_ptr, b.cols,
                               b.diag, static_cast<const float*>(values), static_cast<const float*>(lu), piv, lower, r,
                               z, gate);
        else if (storage == kStoreBF16)
            hipLaunchKernelGGL((k_sgs_group<bf16, float, float>), g, blk, 0, s, rows, count, b.row_ptr, b.cols, b.diag,
                               static_cast<const bf16*>(values), static_cast<const float*>(lu), piv, lower, r, z, gate);
        else
            hipLaunchKernelGGL((k_sgs_group<double, float, float>), g, blk, 0, s, rows, count, b.row_ptr, b.cols,
                               b.diag, static_cast<const double*>(values), static_cast<const float*>(lu), piv, lower, r,
                               z, gate);
    }
}
template void launch_sgs_group<double>(const int*, int, const BsrArgs&, const void*, int, const void*,
                                       const std::uint8_t*, const std::uint8_t*, const double*, double*, hipStream_t,
                                       const double*);
template void launch_sgs_group<float>(const int*, int, const BsrArgs&, const void*, int, const void*,
                                      const std::uint8_t*, const std::uint8_t*, const float*, float*, hipStream_t,
                                      const double*);

template <class T>
void launch_update_p(long long n, double beta, double nbo, const T* r, const T* v, T* p, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_update_p<T>, grid_for(n), dim3(kBlock), 0, s, n, beta, nbo, r, v, p);
}
template void launch_update_p<double>(long long, double, double, const double*, const double*, double*, hipStream_t);
template void launch_update_p<float>(long long, double, double, const float*, const float*, float*, hipStream_t);

template <class T>
void launch_axpy_copy(long long n, const T* x, double a, const T* z, T* y, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_axpy_copy<T>, grid_for(n), dim3(kBlock), 0, s, n, x, a, z, y);
}
template void launch_axpy_copy<double>(long long, const double*, double, const double*, double*, hipStream_t);
template void launch_axpy_copy<float>(long long, const float*, double, const float*, float*, hipStream_t);

template <class T>
void launch_update_xr(long long n, double alpha, double omega, const T* ph, const T* sh, const T* sv, const T* t, T* x,
                      T* r, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_update_xr<T>, grid_for(n), dim3(kBlock), 0, s, n, alpha, omega, ph, sh, sv, t, x, r);
}
template void launch_update_xr<double>(long long, double, double, const double*, const double*, const double*,
                                       const double*, double*, double*, hipStream_t);
template void launch_update_xr<float>(long long, double, double, const float*, const float*, const float*,
                                      const float*, float*, float*, hipStream_t);

void launch_cast_f64_to_f32(long long n, const double* src, float* dst, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL((k_cast<float, double>), grid_for(n), dim3(kBlock), 0, s, n, src, dst);
}

void launch_cast_f32_to_f64(long long n, const float* src, double* dst, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL((k_cast<double, float>), grid_for(n), dim3(kBlock), 0, s, n, src, dst);
}

void launch_add_f32(long long n, const float* e, double* x, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_add_f32, grid_for(n), dim3(kBlock), 0, s, n, e, x);
}

int reduction_blocks(long long n) {
    long long b = (n + kBlock - 1) / kBlock;
    if (b > kReduceBlocks) b = kReduceBlocks;
    if (b < 1) b = 1;
    return static_cast<int>(b);
}

template <class T>
void launch_dots(long long n, int ndot, const DotPairsT<T>& pr, double* partial, double* out, hipStream_t s,
                 const double* gate) {
    const int nb = reduction_blocks(n);
    hipLaunchKernelGGL(k_dots<T>, dim3(nb), dim3(kBlock), 0, s, n, ndot, pr, partial, gate);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, s, nb, ndot, partial, out, gate);
}

template void launch_dots<double>(long long, int, const DotPairsT<double>&, double*, double*, hipStream_t,
                                  const double*);
template void launch_dots<float>(long long, int, const DotPairsT<float>&, double*, double*, hipStream_t,
                                 const double*);

template <class T>
void launch_dots_step(long long n, int ndot, const DotPairsT<T>& pr, double* partial, double* out, double* ks,
                      int step, hipStream_t s) {
    const int nb = reduction_blocks(n);
    hipLaunchKernelGGL(k_dots<T>, dim3(nb), dim3(kBlock), 0, s, n, ndot, pr, partial, ks);
    if (step == kStepAlpha)
        hipLaunchKernelGGL(k_reduce_step<kStepAlpha>, dim3(1), dim3(kBlock), 0, s, nb, ndot, partial, out, ks);
    else if (step == kStepOmega)
        hipLaunchKernelGGL(k_reduce_step<kStepOmega>, dim3(1), dim3(kBlock), 0, s, nb, ndot, partial, out, ks);
    else
        hipLaunchKernelGGL(k_reduce_step<kStepPost>, dim3(1), dim3(kBlock), 0, s, nb, ndot, partial, out, ks);
}
template void launch_dots_step<double>(long long, int, const DotPairsT<double>&, double*, double*, double*, int,
                                       hipStream_t);
template void launch_dots_step<float>(long long, int, const DotPairsT<float>&, double*, double*, double*, int,
                                      hipStream_t);

void launch_bicg_init(double* ks, const double* dots, int x_zero, int fallback, double abs_tol, double rel_tol,
                      int max_iter, hipStream_t s) {
    hipLaunchKernelGGL(k_bicg_init, dim3(1), dim3(1), 0, s, ks, dots, x_zero, fallback, abs_tol, rel_tol, max_iter);
}

template <class T>
void launch_bicg_reject(long long n, const double* ks, const T* b, T* x, T* r, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_bicg_reject<T>, grid_for(n), dim3(kBlock), 0, s, n, ks, b, x, r);
}
template void launch_bicg_reject<double>(long long, const double*, const double*, double*, double*, hipStream_t);
template void launch_bicg_reject<float>(long long, const double*, const float*, float*, float*, hipStream_t);

void launch_bicg_post(double* ks, hipStream_t s) { hipLaunchKernelGGL(k_bicg_post, dim3(1), dim3(1), 0, s, ks); }

template <class T>
void launch_update_p_g(long long n, const double* ks, const T* r, const T* v, T* p, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_update_p_g<T>, grid_for(n), dim3(kBlock), 0, s, n, ks, r, v, p);
}
template void launch_update_p_g<double>(long long, const double*, const double*, const double*, double*, hipStream_t);
template void launch_update_p_g<float>(long long, const double*, const float*, const float*, float*, hipStream_t);

template <class TV>
void launch_vec_sgs_first(int op, int nrows, const double* ks, const TV* r, const TV* v, TV* y,
                          const std::uint8_t* first, int storage, const void* lu, const std::uint8_t* piv, TV* z,
                          hipStream_t s) {
    if (nrows <= 0) return;
    const dim3 g = grid_for((static_cast<long long>(nrows) + kSgsRowsPerWave - 1) / kSgsRowsPerWave * 64), blk(kBlock);
    if constexpr (sizeof(TV) == sizeof(double)) {
        if (storage != kStoreF64) throw_bad_combination("FP64 SGS needs FP64 blocks");
        const double* L = static_cast<const double*>(lu);
        if (op == 0)
            hipLaunchKernelGGL((k_vec_sgs_first<0, double, double>), g, blk, 0, s, nrows, ks, r, v, y, first, L, piv, z);
        else
            hipLaunchKernelGGL((k_vec_sgs_first<1, double, double>), g, blk, 0, s, nrows, ks, r, v, y, first, L, piv, z);
    } else {
        const float* L = static_cast<const float*>(lu);  // FP32 factors with every FP32-vector storage
        if (op == 0)
            hipLaunchKernelGGL((k_vec_sgs_first<0, float, float>), g, blk, 0, s, nrows, ks, r, v, y, first, L, piv, z);
        else
            hipLaunchKernelGGL((k_vec_sgs_first<1, float, float>), g, blk, 0, s, nrows, ks, r, v, y, first, L, piv, z);
    }
}
template void launch_vec_sgs_first<double>(int, int, const double*, const double*, const double*, double*,
                                           const std::uint8_t*, int, const void*, const std::uint8_t*, double*,
                                           hipStream_t);
template void launch_vec_sgs_first<float>(int, int, const double*, const float*, const float*, float*,
                                          const std::uint8_t*, int, const void*, const std::uint8_t*, float*,
                                          hipStream_t);

template <class TV>
void launch_xr_next(int nrows, const double* ks_in, double* ks_out, TV* ph, const TV* sh, const TV* sv, const TV* t,
                    TV* x, TV* r, const TV* v, TV* p, const std::uint8_t* first, int storage, const void* lu,
                    const std::uint8_t* piv, hipStream_t s) {
    if (nrows <= 0) return;
    const dim3 g = grid_for((static_cast<long long>(nrows) + kSgsRowsPerWave - 1) / kSgsRowsPerWave * 64), blk(kBlock);
    if constexpr (sizeof(TV) == sizeof(double)) {
        if (storage != kStoreF64) throw_bad_combination("FP64 SGS needs FP64 blocks");
        hipLaunchKernelGGL((k_xr_next<double, double>), g, blk, 0, s, nrows, ks_in, ks_out, ph, sh, sv, t, x, r, v, p,
                           first, static_cast<const double*>(lu), piv);
    } else {
        hipLaunchKernelGGL((k_xr_next<float, float>), g, blk, 0, s, nrows, ks_in, ks_out, ph, sh, sv, t, x, r, v, p,
                           first, static_cast<const float*>(lu), piv);
    }
}
template void launch_xr_next<double>(int, const double*, double*, double*, const double*, const double*,
                                     const double*, double*, double*, const double*, double*, const std::uint8_t*,
                                     int, const void*, const std::uint8_t*, hipStream_t);
template void launch_xr_next<float>(int, const double*, double*, float*, const float*, const float*, const float*,
                                    float*, float*, const float*, float*, const std::uint8_t*, int, const void*,
                                    const std::uint8_t*, hipStream_t);

template <class T>
void launch_axpy_alpha_g(long long n, const double* ks, const T* r, const T* v, T* sv, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_axpy_alpha_g<T>, grid_for(n), dim3(kBlock), 0, s, n, ks, r, v, sv);
}
template void launch_axpy_alpha_g<double>(long long, const double*, const double*, const double*, double*, hipStream_t);
template void launch_axpy_alpha_g<float>(long long, const double*, const float*, const float*, float*, hipStream_t);

template <class T>
void launch_update_xr_g(long long n, const double* ks, const T* ph, const T* sh, const T* sv, const T* t, T* x, T* r,
                        hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_update_xr_g<T>, grid_for(n), dim3(kBlock), 0, s, n, ks, ph, sh, sv, t, x, r);
}
template void launch_update_xr_g<double>(long long, const double*, const double*, const double*, const double*,
                                         const double*, double*, double*, hipStream_t);
template void launch_update_xr_g<float>(long long, const double*, const float*, const float*, const float*,
                                        const float*, float*, float*, hipStream_t);

void launch_res_sq(int nc, int ld, const double* res, double* partial, double* out, hipStream_t s) {
    const int nb = reduction_blocks(nc);
    hipLaunchKernelGGL(k_res_sq, dim3(nb), dim3(kBlock), 0, s, nc, ld, res, partial);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, s, nb, 5, partial, out, nullptr);
}

namespace {

constexpr int kMaxDevices = 64;

// Device address of `scratch` on the current device, looked up on the first
// call with that device current and kept per device (each GPU holds its own
// copy of the array at its own address).
hipError_t noalloc_scratch(double** out) {
    static std::mutex mu;
    static double* addr[kMaxDevices] = {};
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[dev] = static_cast<double*>(p);
    }
    *out = addr[dev];
    return hipSuccess;
}

}  // namespace

hipError_t residual_norms_noalloc(int nc, int ld, const double* res, double* l2, hipStream_t s) {
    double* partial = nullptr;
    hipError_t e = noalloc_scratch(&partial);
    if (e != hipSuccess) return e;
    const int nb = reduction_blocks(nc);
    const double scale = 1.0 / static_cast<double>(nc > 1 ? nc : 1);
    hipLaunchKernelGGL(k_res_sq, dim3(nb), dim3(kBlock), 0, s, nc, ld, res, partial);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_norms_finish, dim3(1), dim3(kBlock), 0, s, nb, partial, scale, l2);
    return hipGetLastError();
}

void launch_rhs(int nc, int ld, const double* res, double* b, hipStream_t s) {
    if (nc > 0) hipLaunchKernelGGL(k_rhs, grid_for(nc), dim3(kBlock), 0, s, nc, ld, res, b);
}

void launch_state_update(int nc, int ld, const double* du, double* U, hipStream_t s) {
    if (nc > 0) hipLaunchKernelGGL(k_state_update, grid_for(nc), dim3(kBlock), 0, s, nc, ld, du, U);
}

}  // namespace cfdg
