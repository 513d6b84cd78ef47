// Block-sparse (5x5 BSR) linear-algebra kernels for the implicit step (gfx950).
//
//   k_lu_factor   diagonal-block LU with partial pivoting (bsr_helpers.hpp:278-321)
//   k_spmv        y = A x, five lanes per block row (bsr_helpers.hpp:139-203)
//   k_sgs_group   one Gauss-Seidel row update for every row of one row group
//                 (bsr_helpers.hpp:390-435 for a single row); a sweep launches
//                 the groups in order (colours, or dependency levels)
//   vector ops    the BiCGSTAB updates of bicgstab.cpp:97-177, fused
//   reductions    dot products: per-block partial sums in a fixed order, then
//                 one block adds the partials in a fixed order -> the result
//                 is identical from run to run (no atomics)
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"
#include "kernels_linalg.hpp"

namespace cfdg {

namespace {

constexpr int kBlock = 256;

inline dim3 grid_for(long long n) { return dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)); }

__global__ void __launch_bounds__(kBlock) k_lu_factor(int n, const int* diag, const double* values, double* lu,
                                                      std::uint8_t* piv, int* singular) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // fully unrolled so that the 25 entries stay in registers: the row swap is
    // a select over the (compile-time) candidate rows
    double a[25];
    const double* src = values + static_cast<long long>(diag[i]) * 25;
#pragma unroll
    for (int k = 0; k < 25; ++k) a[k] = src[k];
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
    double* dst = lu + static_cast<long long>(i) * 25;
#pragma unroll
    for (int k = 0; k < 25; ++k) dst[k] = a[k];
#pragma unroll
    for (int k = 0; k < 5; ++k) piv[static_cast<long long>(i) * 5 + k] = static_cast<std::uint8_t>(p[k]);
}

// y = A x with five lanes per block row: lane ii forms
// acc_ii = sum_k (sum_jj a_k[ii][jj] x_col(k)[jj]) in column order.
__global__ void __launch_bounds__(kBlock) k_spmv(int n, const int* row_ptr, const int* cols, const double* values,
                                                 const double* x, double* y) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long row = t / 5;
    const int ii = static_cast<int>(t - row * 5);
    if (row >= n) return;
    double acc = 0.0;
    const int kb = row_ptr[row], ke = row_ptr[row + 1];
    for (int k = kb; k < ke; ++k) {
        const double* a = values + static_cast<long long>(k) * 25 + ii * 5;
        const double* xb = x + static_cast<long long>(cols[k]) * 5;
        double s = 0.0;
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
__global__ void __launch_bounds__(kBlock) k_sgs_group(const int* rows, int count, const int* row_ptr, const int* cols,
                                                      const int* diag, const double* values, const double* lu,
                                                      const std::uint8_t* piv, const double* r, double* z) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int i = rows[t];
    const int d = diag[i];
    double s[5];
    for (int ii = 0; ii < 5; ++ii) s[ii] = r[static_cast<long long>(i) * 5 + ii];
    const int kb = row_ptr[i], ke = row_ptr[i + 1];
    for (int k = kb; k < ke; ++k) {
        if (k == d) continue;
        const double* a = values + static_cast<long long>(k) * 25;
        const double* zb = z + static_cast<long long>(cols[k]) * 5;
        const double z0 = zb[0], z1 = zb[1], z2 = zb[2], z3 = zb[3], z4 = zb[4];
        for (int ii = 0; ii < 5; ++ii) {
            double tt = 0.0;
            tt += a[ii * 5 + 0] * z0;
            tt += a[ii * 5 + 1] * z1;
            tt += a[ii * 5 + 2] * z2;
            tt += a[ii * 5 + 3] * z3;
            tt += a[ii * 5 + 4] * z4;
            s[ii] -= tt;
        }
    }
    // lu_solve_block<5> (bsr_helpers.hpp:325-353)
    const double* L = lu + static_cast<long long>(i) * 25;
    const std::uint8_t* p = piv + static_cast<long long>(i) * 5;
    for (int k = 0; k < 5; ++k) {
        const int pk = p[k];
        if (pk != k) {
            double tmp = s[k];
            // s[pk] with a compile-time-friendly select
            double sp = s[0];
            for (int q = 1; q < 5; ++q) sp = (q == pk) ? s[q] : sp;
            s[k] = sp;
            for (int q = 0; q < 5; ++q) s[q] = (q == pk) ? tmp : s[q];
        }
    }
    for (int a = 1; a < 5; ++a) {
        double acc = s[a];
        for (int b = 0; b < a; ++b) acc -= L[a * 5 + b] * s[b];
        s[a] = acc;
    }
    for (int a = 4; a >= 0; --a) {
        double acc = s[a];
        for (int b = a + 1; b < 5; ++b) acc -= L[a * 5 + b] * s[b];
        s[a] = acc / L[a * 5 + a];
    }
    for (int ii = 0; ii < 5; ++ii) z[static_cast<long long>(i) * 5 + ii] = s[ii];
}

// p = ((p * beta) + (-beta * omega) v) + r           (bicgstab.cpp:104-109)
__global__ void __launch_bounds__(kBlock) k_update_p(long long n, double beta, double nbo, const double* r,
                                                     const double* v, double* p) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double t = p[i] * beta;
    t += nbo * v[i];
    t += 1.0 * r[i];
    p[i] = t;
}

// y = x + a z
__global__ void __launch_bounds__(kBlock) k_axpy_copy(long long n, const double* x, double a, const double* z,
                                                      double* y) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = x[i] + a * z[i];
}

// x = (x + alpha ph) + omega sh;  r = s + (-omega) t   (bicgstab.cpp:145-149)
__global__ void __launch_bounds__(kBlock) k_update_xr(long long n, double alpha, double omega, const double* ph,
                                                      const double* sh, const double* s, const double* t, double* x,
                                                      double* r) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double xv = x[i];
    xv += alpha * ph[i];
    xv += omega * sh[i];
    x[i] = xv;
    double rv = s[i];
    rv += (-omega) * t[i];
    r[i] = rv;
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
__global__ void __launch_bounds__(kBlock) k_dots(long long n, int ndot, DotPairs pr, double* partial) {
    __shared__ double lds[kBlock / 64];
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        for (int j = 0; j < 5; ++j)
            if (j < ndot) acc[j] += pr.a[j][i] * pr.b[j][i];
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
                                                            double* out) {
    __shared__ double lds[kBlock / 64];
    for (int j = 0; j < ncomp; ++j) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += kBlock) s += partial[b * 5 + j];
        s = block_sum(s, lds);
        if (threadIdx.x == 0) out[j] = s;
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

void launch_lu_factor(int n, const int* diag, const double* values, double* lu, std::uint8_t* piv, int* singular,
                      hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_lu_factor, grid_for(n), dim3(kBlock), 0, s, n, diag, values, lu, piv, singular);
}

void launch_spmv(const BsrArgs& b, const double* values, const double* x, double* y, hipStream_t s) {
    if (b.n > 0)
        hipLaunchKernelGGL(k_spmv, grid_for(5LL * b.n), dim3(kBlock), 0, s, b.n, b.row_ptr, b.cols, values, x, y);
}

void launch_sgs_group(const int* rows, int count, const BsrArgs& b, const double* values, const double* lu,
                      const std::uint8_t* piv, const double* r, double* z, hipStream_t s) {
    if (count > 0)
        hipLaunchKernelGGL(k_sgs_group, grid_for(count), dim3(kBlock), 0, s, rows, count, b.row_ptr, b.cols, b.diag,
                           values, lu, piv, r, z);
}

void launch_update_p(long long n, double beta, double nbo, const double* r, const double* v, double* p,
                     hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_update_p, grid_for(n), dim3(kBlock), 0, s, n, beta, nbo, r, v, p);
}

void launch_axpy_copy(long long n, const double* x, double a, const double* z, double* y, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_axpy_copy, grid_for(n), dim3(kBlock), 0, s, n, x, a, z, y);
}

void launch_update_xr(long long n, double alpha, double omega, const double* ph, const double* sh, const double* sv,
                      const double* t, double* x, double* r, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_update_xr, grid_for(n), dim3(kBlock), 0, s, n, alpha, omega, ph, sh, sv, t, x, r);
}

int reduction_blocks(long long n) {
    long long b = (n + kBlock - 1) / kBlock;
    if (b > kReduceBlocks) b = kReduceBlocks;
    if (b < 1) b = 1;
    return static_cast<int>(b);
}

void launch_dots(long long n, int ndot, const DotPairs& pr, double* partial, double* out, hipStream_t s) {
    const int nb = reduction_blocks(n);
    hipLaunchKernelGGL(k_dots, dim3(nb), dim3(kBlock), 0, s, n, ndot, pr, partial);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, s, nb, ndot, partial, out);
}

void launch_res_sq(int nc, int ld, const double* res, double* partial, double* out, hipStream_t s) {
    const int nb = reduction_blocks(nc);
    hipLaunchKernelGGL(k_res_sq, dim3(nb), dim3(kBlock), 0, s, nc, ld, res, partial);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, s, nb, 5, partial, out);
}

void launch_rhs(int nc, int ld, const double* res, double* b, hipStream_t s) {
    if (nc > 0) hipLaunchKernelGGL(k_rhs, grid_for(nc), dim3(kBlock), 0, s, nc, ld, res, b);
}

void launch_state_update(int nc, int ld, const double* du, double* U, hipStream_t s) {
    if (nc > 0) hipLaunchKernelGGL(k_state_update, grid_for(nc), dim3(kBlock), 0, s, nc, ld, du, U);
}

}  // namespace cfdg
