// Kernels of the learned warm start (gnn.hpp): BF16 MFMA node products,
// FP32 edge layers, loss, backward and Adam (gfx950).
//
// MFMA: v_mfma_f32_16x16x32_bf16, D(16x16) += A(16x32) B(32x16) per wave.
// Operand lanes (l = 0..63, i = 0..7, r = 0..3):
//   A[l % 16][8 (l / 16) + i],  B[8 (l / 16) + i][l % 16],  D[4 (l / 16) + r][l % 16].
#include <hip/hip_runtime.h>

#include <cmath>

#include "gnn.hpp"
#include "kernels.hpp"

namespace cfdg {
namespace gnn {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;

__device__ __forceinline__ short bf16_bits_from_float(float f) { return static_cast<short>(bf16_from_float(f).u); }

__device__ __forceinline__ bf16x8_t load_bf16x8(const bf16* p) {
    return *reinterpret_cast<const bf16x8_t*>(p);
}

// C[n x M] = A[n x k] W[M x k]^T; one wave per 16 rows, the block covers 64 rows.
template <int M>
__global__ void __launch_bounds__(kBlock) k_gemm_nt(int n, int k, const bf16* A, const bf16* W, float* C) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int row0 = (blockIdx.x * kWaves + wave) * 16;
    if (row0 >= n) return;
    constexpr int T = M / 16;
    f32x4_t acc[T];
    for (int t = 0; t < T; ++t) acc[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    const int ar = row0 + (lane & 15);
    const int kq = 8 * (lane >> 4);
    for (int kk = 0; kk < k; kk += 32) {
        bf16x8_t a = {0, 0, 0, 0, 0, 0, 0, 0};
        if (ar < n) a = load_bf16x8(A + static_cast<long long>(ar) * k + kk + kq);
        for (int t = 0; t < T; ++t) {
            const bf16x8_t b = load_bf16x8(W + static_cast<long long>(t * 16 + (lane & 15)) * k + kk + kq);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
        }
    }
    const int col = lane & 15;
    for (int t = 0; t < T; ++t) {
        for (int r = 0; r < 4; ++r) {
            const int row = row0 + 4 * (lane >> 4) + r;
            if (row < n) C[static_cast<long long>(row) * M + t * 16 + col] = acc[t][r];
        }
    }
}

// partial[b][m][k] = sum over rows of chunk b of dY[row][m] A[row][k]: the
// contracted MFMA dimension runs over rows (32 per instruction).  Each wave
// owns output tiles (tm, tk) round robin.
__global__ void __launch_bounds__(kBlock) k_gemm_tn(int n, int m, int k, int chunk, const float* dY, const bf16* A,
                                                    float* partial) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int r0 = blockIdx.x * chunk;
    const int r1 = min(n, r0 + chunk);
    const int tiles_m = m / 16, tiles_k = k / 16;
    float* out = partial + static_cast<long long>(blockIdx.x) * m * k;
    for (int tile = wave; tile < tiles_m * tiles_k; tile += kWaves) {
        const int tm = tile / tiles_k, tk = tile % tiles_k;
        f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
        for (int rr = r0; rr < r1; rr += 32) {
            bf16x8_t a, b;
            for (int i = 0; i < 8; ++i) {
                const int row = rr + 8 * (lane >> 4) + i;
                const bool in = row < r1;
                const float dy = in ? dY[static_cast<long long>(row) * m + tm * 16 + (lane & 15)] : 0.f;
                a[i] = bf16_bits_from_float(dy);
                b[i] = in ? static_cast<short>(A[static_cast<long long>(row) * k + tk * 16 + (lane & 15)].u) : short(0);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        for (int r = 0; r < 4; ++r)
            out[(tm * 16 + 4 * (lane >> 4) + r) * k + tk * 16 + (lane & 15)] = acc[r];
    }
}

__global__ void __launch_bounds__(kBlock) k_reduce_partials(int nblocks, int len, const float* partial, float* out,
                                                            int accumulate) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    float s = 0.f;
    for (int b = 0; b < nblocks; ++b) s += partial[static_cast<long long>(b) * len + i];
    out[i] = accumulate ? out[i] + s : s;
}

// One wave per node, lane u = hidden unit.  Optional output layer
// y = Wo h + bo by a fixed wave reduction.
__global__ void __launch_bounds__(kBlock) k_edge_forward(Graph g, const float* pq, const float* Gm, const float* b,
                                                         bf16* h_bf16, float* h_f32, const float* Wo, const float* bo,
                                                         float* y) {
    const int u = threadIdx.x & 63;
    const int node = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (node >= g.n) return;
    const float P = pq[static_cast<long long>(node) * kPQ + u];
    const float bu = b[u];
    const float g0 = Gm[u * kG + 0], g1 = Gm[u * kG + 1], g2 = Gm[u * kG + 2], g3 = Gm[u * kG + 3];
    const int e0 = g.adj_off[node], e1 = g.adj_off[node + 1];
    float acc = 0.f;
    for (int e = e0; e < e1; ++e) {
        const int j = g.adj[e];
        const float* ge = g.eg + static_cast<long long>(e) * kG;
        float z = P + pq[static_cast<long long>(j) * kPQ + kH + u];
        z += g0 * ge[0] + g1 * ge[1] + g2 * ge[2] + g3 * ge[3] + bu;
        acc += fmaxf(z, 0.f);
    }
    const float h = (e1 > e0) ? acc / static_cast<float>(e1 - e0) : 0.f;
    if (h_bf16) h_bf16[static_cast<long long>(node) * kH + u] = bf16_from_float(h);
    if (h_f32) h_f32[static_cast<long long>(node) * kH + u] = h;
    if (Wo) {
        for (int o = 0; o < kO; ++o) {
            float v = Wo[o * kH + u] * h;
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (u == 0) y[static_cast<long long>(node) * kO + o] = v + bo[o];
        }
    }
}

// Loss (y - t)^2 / (n kO) and the output-layer backward.  Grid-stride over
// nodes (fixed grid -> fixed order); per block partial[b] = {dWo (kO x kH),
// dbo (kO), loss}.
constexpr int kOutPartial = kO * kH + kO + 1;
constexpr int kOutBlocks = 256;

__global__ void __launch_bounds__(kBlock) k_out_backward(int n, const float* y, const float* t, const float* h2,
                                                         const float* Wo, float* dh2, float* partial) {
    __shared__ float red[kWaves][kOutPartial];
    const int u = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const float scale = 2.0f / (static_cast<float>(n) * kO);
    float dW[kO] = {0.f, 0.f, 0.f, 0.f, 0.f};
    float db[kO] = {0.f, 0.f, 0.f, 0.f, 0.f};
    float loss = 0.f;
    for (int node = blockIdx.x * kWaves + wave; node < n; node += gridDim.x * kWaves) {
        float dy[kO];
        for (int o = 0; o < kO; ++o) {
            const float d = y[static_cast<long long>(node) * kO + o] - t[static_cast<long long>(node) * kO + o];
            dy[o] = scale * d;
            if (u == 0) loss += d * d;
        }
        const float h = h2[static_cast<long long>(node) * kH + u];
        float dh = 0.f;
        for (int o = 0; o < kO; ++o) {
            dh += Wo[o * kH + u] * dy[o];
            dW[o] += dy[o] * h;
            db[o] += dy[o];
        }
        dh2[static_cast<long long>(node) * kH + u] = dh;
    }
    for (int o = 0; o < kO; ++o) red[wave][o * kH + u] = dW[o];
    if (u < kO) red[wave][kO * kH + u] = db[u];
    if (u == 0) red[wave][kO * kH + kO] = loss / (static_cast<float>(n) * kO);
    __syncthreads();
    for (int i = threadIdx.x; i < kOutPartial; i += kBlock) {
        float s = 0.f;
        for (int w = 0; w < kWaves; ++w) s += red[w][i];
        partial[static_cast<long long>(blockIdx.x) * kOutPartial + i] = s;
    }
}

// Edge layer backward, one wave per node (grid-stride, fixed grid):
//   dz_e = (dh_i / deg_i) [z_e > 0];  dP_i = sum_{e in adj(i)} dz_e;
//   dQ_j = sum over edges e = (i -> j) of dz_e, gathered at j through the
//   reverse edges;  dG += dz_e g_e^T;  db += dz_e.
constexpr int kEdgePartial = kH * kG + kH;
constexpr int kEdgeBlocks = 512;

__global__ void __launch_bounds__(kBlock) k_edge_backward(Graph g, const float* pq, const float* Gm, const float* b,
                                                          const float* dh, float* dpq, float* partial) {
    __shared__ float red[kWaves][kEdgePartial];
    const int u = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const float bu = b[u];
    const float g0 = Gm[u * kG + 0], g1 = Gm[u * kG + 1], g2 = Gm[u * kG + 2], g3 = Gm[u * kG + 3];
    float dG[kG] = {0.f, 0.f, 0.f, 0.f};
    float dbu = 0.f;
    for (int node = blockIdx.x * kWaves + wave; node < g.n; node += gridDim.x * kWaves) {
        const int e0 = g.adj_off[node], e1 = g.adj_off[node + 1];
        // outgoing edges: dP and the parameter gradients
        const float Pi = pq[static_cast<long long>(node) * kPQ + u];
        const float di = (e1 > e0) ? dh[static_cast<long long>(node) * kH + u] / static_cast<float>(e1 - e0) : 0.f;
        float dP = 0.f;
        for (int e = e0; e < e1; ++e) {
            const int j = g.adj[e];
            const float* ge = g.eg + static_cast<long long>(e) * kG;
            float z = Pi + pq[static_cast<long long>(j) * kPQ + kH + u];
            z += g0 * ge[0] + g1 * ge[1] + g2 * ge[2] + g3 * ge[3] + bu;
            const float dz = (z > 0.f) ? di : 0.f;
            dP += dz;
            dG[0] += dz * ge[0];
            dG[1] += dz * ge[1];
            dG[2] += dz * ge[2];
            dG[3] += dz * ge[3];
            dbu += dz;
        }
        // incoming edges (i -> node): dQ
        const float Qn = pq[static_cast<long long>(node) * kPQ + kH + u];
        float dQ = 0.f;
        for (int e = e0; e < e1; ++e) {
            const int i = g.adj[e];
            const int er = g.rev[e];
            const int ie0 = g.adj_off[i], ie1 = g.adj_off[i + 1];
            const float* ge = g.eg + static_cast<long long>(er) * kG;
            float z = pq[static_cast<long long>(i) * kPQ + u] + Qn;
            z += g0 * ge[0] + g1 * ge[1] + g2 * ge[2] + g3 * ge[3] + bu;
            const float dii = dh[static_cast<long long>(i) * kH + u] / static_cast<float>(ie1 - ie0);
            dQ += (z > 0.f) ? dii : 0.f;
        }
        dpq[static_cast<long long>(node) * kPQ + u] = dP;
        dpq[static_cast<long long>(node) * kPQ + kH + u] = dQ;
    }
    for (int q = 0; q < kG; ++q) red[wave][u * kG + q] = dG[q];
    red[wave][kH * kG + u] = dbu;
    __syncthreads();
    for (int i = threadIdx.x; i < kEdgePartial; i += kBlock) {
        float s = 0.f;
        for (int w = 0; w < kWaves; ++w) s += red[w][i];
        partial[static_cast<long long>(blockIdx.x) * kEdgePartial + i] = s;
    }
}

__global__ void __launch_bounds__(kBlock) k_to_bf16(int n, const float* src, bf16* dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = bf16_from_float(src[i]);
}

// dst[c][r] = src[r][c] (src rows x cols), as bf16
__global__ void __launch_bounds__(kBlock) k_transpose_to_bf16(int rows, int cols, const float* src, bf16* dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * cols) return;
    const int r = i / cols, c = i % cols;
    dst[c * rows + r] = bf16_from_float(src[i]);
}

__global__ void __launch_bounds__(kBlock) k_adam(int n, float* p, const float* grad, float* m, float* v, float lr,
                                                 float beta1, float beta2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gr = grad[i];
    const float mi = beta1 * m[i] + (1.f - beta1) * gr;
    const float vi = beta2 * v[i] + (1.f - beta2) * gr * gr;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void __launch_bounds__(kBlock) k_build_graph(int nc, int ninner, const int* cf_off, const int* cf,
                                                        const double* fp, long long fld, const int2* fcells,
                                                        const double* cs, long long ld, int* adj, int* rev, float* eg) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const double vol = cs[plane_at(kRVol, ld, c)];
    const double v23 = cbrt(vol) * cbrt(vol);
    for (int e = cf_off[c]; e < cf_off[c + 1]; ++e) {
        const int f = cf[e];
        const FaceRef fr{fp, fld, f};
        const int cells[2] = {fcells[f].x, fcells[f].y};
        const bool inner = f < ninner;
        const int j = !inner ? c : (cells[0] == c ? cells[1] : cells[0]);
        const double sg = (cells[0] == c) ? 1.0 : -1.0;
        adj[e] = j;
        int r = e;
        if (inner) {
            for (int k = cf_off[j]; k < cf_off[j + 1]; ++k)
                if (cf[k] == f) r = k;
        }
        rev[e] = r;
        float* ge = eg + static_cast<long long>(e) * kG;
        ge[0] = static_cast<float>(sg * fr[kFN]);
        ge[1] = static_cast<float>(sg * fr[kFN + 1]);
        ge[2] = static_cast<float>(sg * fr[kFN + 2]);
        ge[3] = static_cast<float>(fr[kFA] / v23);
    }
}

// Residual scaled by the cell's spectral radius, -R / lambda (the update of a
// unit-CFL explicit step), in units of the conservative freestream scales.
// The implicit update (alpha I + J)^-1 (-R) is of this size at any CFL.
__device__ __forceinline__ void explicit_update(const double* res, const double* lam, int nc, int c, const Ref& ref,
                                                double e[5]) {
    const double f = 1.0 / fmax(lam[c], 1.0e-300);
    for (int v = 0; v < 5; ++v) e[v] = -res[static_cast<long long>(v) * nc + c] * f / ref.s[v];
}

constexpr int kNormBlocks = 512;

__global__ void __launch_bounds__(kBlock) k_explicit_norm(int nc, const double* res, const double* lam, Ref ref,
                                                          double* partial) {
    __shared__ double red[kBlock];
    double acc = 0.0;
    for (int c = blockIdx.x * kBlock + threadIdx.x; c < nc; c += gridDim.x * kBlock) {
        double e[5];
        explicit_update(res, lam, nc, c, ref, e);
        for (int v = 0; v < 5; ++v) acc += e[v] * e[v];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_explicit_norm_final(int nb, const double* partial, double* out) {
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += partial[b];
    out[0] = s;
}

__device__ __forceinline__ double rms_of(const double* rnorm2, int nc) {
    const double r = sqrt(rnorm2[0] / (5.0 * static_cast<double>(nc > 0 ? nc : 1)));
    return r > 1.0e-300 ? r : 1.0e-300;
}

__global__ void __launch_bounds__(kBlock) k_features(int nc, const double* cs, long long ld, const double* res,
                                                     const double* lam, const double* dt, Ref ref, const double* rnorm2,
                                                     const double* du_prev, int have_prev, bf16* x) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const double rn = rms_of(rnorm2, nc);
    const CellRef r{cs, ld, c};
    double e[5];
    explicit_update(res, lam, nc, c, ref, e);
    const double h = cbrt(r[kRVol]);
    float f[kF];
    for (int k = 0; k < kF; ++k) f[k] = 0.f;
    f[0] = static_cast<float>(r[kRQ] / ref.p - 1.0);
    f[1] = static_cast<float>(r[kRQ + 1] / ref.speed);
    f[2] = static_cast<float>(r[kRQ + 2] / ref.speed);
    f[3] = static_cast<float>(r[kRQ + 3] / ref.speed);
    f[4] = static_cast<float>(r[kRQ + 4] / ref.T - 1.0);
    for (int v = 0; v < 5; ++v) f[5 + v] = static_cast<float>(e[v] / rn);
    f[10] = static_cast<float>(log10(rn) / 8.0);
    f[11] = static_cast<float>(ref.mach);
    f[12] = static_cast<float>(log(h / ref.h_mean));
    f[13] = static_cast<float>(log10(dt[c] * (ref.speed + ref.a) / h) / 3.0);
    f[14] = 1.f;
    if (have_prev) {
        for (int v = 0; v < 5; ++v)
            f[15 + v] = static_cast<float>(du_prev[static_cast<long long>(c) * 5 + v] / (ref.s[v] * rn));
        const double rp = sqrt(rnorm2[1] / (5.0 * static_cast<double>(nc)));
        f[20] = static_cast<float>(rp > 1.0e-300 ? log(rn / rp) : 0.0);
        f[21] = 1.f;
    }
    bf16* xo = x + static_cast<long long>(c) * kF;
    for (int k = 0; k < kF; ++k) xo[k] = bf16_from_float(f[k]);
}

// The network predicts the change of the update: du_n = du_{n-1} + y s rnorm
// (du_{n-1} = 0 on the first step), a skip connection carried in FP64.
__global__ void __launch_bounds__(kBlock) k_targets(int nc, const double* du, const double* du_prev, int have_prev,
                                                    Ref ref, const double* rnorm2, float* t) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const double rn = rms_of(rnorm2, nc);
    for (int v = 0; v < 5; ++v) {
        const long long i = static_cast<long long>(c) * 5 + v;
        const double d = have_prev ? du[i] - du_prev[i] : du[i];
        t[static_cast<long long>(c) * kO + v] = static_cast<float>(d / (ref.s[v] * rn));
    }
}

__global__ void __launch_bounds__(kBlock) k_prediction(int nc, const float* y, const double* du_prev, int have_prev,
                                                       Ref ref, const double* rnorm2, double* x0) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const double rn = rms_of(rnorm2, nc);
    for (int v = 0; v < 5; ++v) {
        const long long i = static_cast<long long>(c) * 5 + v;
        const double corr = static_cast<double>(y[static_cast<long long>(c) * kO + v]) * ref.s[v] * rn;
        x0[i] = have_prev ? du_prev[i] + corr : corr;
    }
}

inline dim3 grid_for(long long n) { return dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)); }

}  // namespace

void launch_build_graph(int nc, int ninner, const int* cf_off, const int* cf, const double* fp, long long fld,
                        const int2* fcells, const double* cs, long long ld, int* adj, int* rev, float* eg,
                        hipStream_t s) {
    if (nc > 0)
        hipLaunchKernelGGL(k_build_graph, grid_for(nc), dim3(kBlock), 0, s, nc, ninner, cf_off, cf, fp, fld, fcells, cs,
                           ld, adj, rev, eg);
}

void launch_explicit_norm(int nc, const double* res, const double* lam, Ref ref, double* partial, double* out,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_explicit_norm, dim3(kNormBlocks), dim3(kBlock), 0, s, nc, res, lam, ref, partial);
    hipLaunchKernelGGL(k_explicit_norm_final, dim3(1), dim3(64), 0, s, kNormBlocks, partial, out);
}

void launch_features(int nc, const double* cs, long long ld, const double* res, const double* lam, const double* dt,
                     Ref ref, const double* rnorm2, const double* du_prev, int have_prev, bf16* x, hipStream_t s) {
    if (nc > 0)
        hipLaunchKernelGGL(k_features, grid_for(nc), dim3(kBlock), 0, s, nc, cs, ld, res, lam, dt, ref, rnorm2,
                           du_prev, have_prev, x);
}

void launch_targets(int nc, const double* du, const double* du_prev, int have_prev, Ref ref, const double* rnorm2,
                    float* t, hipStream_t s) {
    if (nc > 0)
        hipLaunchKernelGGL(k_targets, grid_for(nc), dim3(kBlock), 0, s, nc, du, du_prev, have_prev, ref, rnorm2, t);
}

void launch_prediction(int nc, const float* y, const double* du_prev, int have_prev, Ref ref, const double* rnorm2,
                       double* x0, hipStream_t s) {
    if (nc > 0)
        hipLaunchKernelGGL(k_prediction, grid_for(nc), dim3(kBlock), 0, s, nc, y, du_prev, have_prev, ref, rnorm2, x0);
}

void launch_gemm_nt(int n, int m, int k, const bf16* A, const bf16* W, float* C, hipStream_t s) {
    if (n <= 0) return;
    const dim3 grid(static_cast<unsigned>((n + 16 * kWaves - 1) / (16 * kWaves)));
    if (m == 128)
        hipLaunchKernelGGL(k_gemm_nt<128>, grid, dim3(kBlock), 0, s, n, k, A, W, C);
    else if (m == 64)
        hipLaunchKernelGGL(k_gemm_nt<64>, grid, dim3(kBlock), 0, s, n, k, A, W, C);
    else
        hipLaunchKernelGGL(k_gemm_nt<16>, grid, dim3(kBlock), 0, s, n, k, A, W, C);
}

int gemm_tn_blocks(int n, int chunk) { return (n + chunk - 1) / chunk; }

void launch_gemm_tn(int n, int m, int k, int chunk, const float* dY, const bf16* A, float* partial, hipStream_t s) {
    const int nb = gemm_tn_blocks(n, chunk);
    if (nb > 0) hipLaunchKernelGGL(k_gemm_tn, dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, s, n, m, k, chunk, dY, A, partial);
}

void launch_reduce_partials(int nblocks, int len, const float* partial, float* out, int accumulate, hipStream_t s) {
    if (len > 0) hipLaunchKernelGGL(k_reduce_partials, grid_for(len), dim3(kBlock), 0, s, nblocks, len, partial, out, accumulate);
}

void launch_edge_forward(const Graph& g, const float* pq, const float* Gm, const float* b, bf16* h_bf16, float* h_f32,
                         const float* Wo, const float* bo, float* y, hipStream_t s) {
    if (g.n <= 0) return;
    const dim3 grid(static_cast<unsigned>((g.n + kWaves - 1) / kWaves));
    hipLaunchKernelGGL(k_edge_forward, grid, dim3(kBlock), 0, s, g, pq, Gm, b, h_bf16, h_f32, Wo, bo, y);
}

int out_backward_blocks(int) { return kOutBlocks; }

void launch_out_backward(int n, const float* y, const float* t, const float* h2, const float* Wo, float* dh2,
                         float* partial, hipStream_t s) {
    hipLaunchKernelGGL(k_out_backward, dim3(kOutBlocks), dim3(kBlock), 0, s, n, y, t, h2, Wo, dh2, partial);
}

int edge_backward_blocks(int) { return kEdgeBlocks; }

void launch_edge_backward(const Graph& g, const float* pq, const float* Gm, const float* b, const float* dh,
                          float* dpq, float* partial, hipStream_t s) {
    hipLaunchKernelGGL(k_edge_backward, dim3(kEdgeBlocks), dim3(kBlock), 0, s, g, pq, Gm, b, dh, dpq, partial);
}

void launch_to_bf16(int n, const float* src, bf16* dst, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_to_bf16, grid_for(n), dim3(kBlock), 0, s, n, src, dst);
}

void launch_transpose_to_bf16(int rows, int cols, const float* src, bf16* dst, hipStream_t s) {
    if (rows * cols > 0)
        hipLaunchKernelGGL(k_transpose_to_bf16, grid_for(static_cast<long long>(rows) * cols), dim3(kBlock), 0, s, rows,
                           cols, src, dst);
}

void launch_adam(int n, float* p, const float* grad, float* m, float* v, float lr, float beta1, float beta2, float eps,
                 int t, hipStream_t s) {
    const float c1 = 1.f - std::pow(beta1, static_cast<float>(t));
    const float c2 = 1.f - std::pow(beta2, static_cast<float>(t));
    if (n > 0) hipLaunchKernelGGL(k_adam, grid_for(n), dim3(kBlock), 0, s, n, p, grad, m, v, lr, beta1, beta2, eps, c1, c2);
}

}  // namespace gnn
}  // namespace cfdg
