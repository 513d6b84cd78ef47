// Learned warm start: a two-layer edge-convolution network over the face
// graph that predicts the Newton update of an implicit step (no counterpart
// in the reference; its linear solver always starts from zero,
// implicit_euler.hpp:81-92).
//
// Graph: nodes = cells (device order), one directed edge per face of each
// cell (adj_off = the cell -> faces offsets): an interior face leads to the
// neighbour, a boundary face is a self loop.  Edge geometry g_e = (n_x, n_y,
// n_z, A / V_i^(2/3)) with the normal pointing away from node i.
//
// Network (F = 32 input features, 24 used; H = 64; G = 4; O = 5):
//   z1_e   = WP1 x_i + WQ1 x_j + G1 g_e + b1          e = (i -> j)
//   h1_i   = mean_{e in adj(i)} relu(z1_e)
//   z2_e   = WP2 h1_i + WQ2 h1_j + G2 g_e + b2
//   h2_i   = mean_{e in adj(i)} relu(z2_e)
//   y_i    = Wo h2_i + bo   (the predicted update is du_prev + y s rnorm)
// (the EdgeConv form relu(A x_i + B (x_j - x_i)) with WP = A - B, WQ = B).
// The node products [WP; WQ] x run on MFMA in BF16 with FP32 accumulation;
// the edge sums, the output layer, the loss and Adam run in FP32.  Every sum
// has a fixed order: results are deterministic.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "precision.hpp"

namespace cfdg {
namespace gnn {

constexpr int kF = 32;   // stored input features (bf16); kFUsed carry data, the rest are zero
constexpr int kFUsed = 24;
constexpr int kH = 64;
constexpr int kG = 4;
constexpr int kO = 5;
constexpr int kPQ = 2 * kH;  // stacked [WP; WQ] outputs

// Parameter layout (FP32, flat): offsets in floats.
constexpr int kOffW1 = 0;                        // [WP1; WQ1]  kPQ x kF (columns >= kFUsed are zero and stay zero)
constexpr int kOffG1 = kOffW1 + kPQ * kF;        // kH x kG
constexpr int kOffB1 = kOffG1 + kH * kG;         // kH
constexpr int kOffW2 = kOffB1 + kH;              // [WP2; WQ2]  kPQ x kH
constexpr int kOffG2 = kOffW2 + kPQ * kH;        // kH x kG
constexpr int kOffB2 = kOffG2 + kH * kG;         // kH
constexpr int kOffWo = kOffB2 + kH;              // kO x kH
constexpr int kOffBo = kOffWo + kO * kH;         // kO
constexpr int kNumParams = kOffBo + kO;

struct Graph {
    int n;              // nodes
    const int* adj_off; // n + 1
    const int* adj;     // neighbour of each directed edge
    const int* rev;     // index of the reverse edge (j -> i) of each edge (i -> j)
    const float* eg;    // kG per edge
};

// Forward buffers of one graph.
struct Acts {
    const bf16* x;   // n x kF
    float* pq1;      // n x kPQ
    bf16* h1;        // n x kH
    float* pq2;      // n x kPQ
    float* h2;       // n x kH
    float* y;        // n x kO
};

// Freestream reference of the feature / target normalisation.
struct Ref {
    double p, T, speed, a, rho, mach;
    double s[5];    // conservative scales: rho, rho U, rho U, rho U, rho E
    double h_mean;  // mean V^(1/3) of the mesh
};

// ---- graph and feature kernels (cells in the device order of the residual) ----
// Directed edges = the faces of each cell in the cell -> faces lists (adj_off = cf_off):
// an interior face links the cell to its neighbour, a boundary face is a self loop.
// eg = (n oriented away from the cell, A / V^(2/3)); rev = the same face seen from the other cell.
// (cell and face planes as in kernels.hpp)
void launch_build_graph(int nc, int ninner, const int* cf_off, const int* cf, const double* fp, long long fld,
                        const int2* fcells, const double* cs, long long ld, int* adj, int* rev, float* eg,
                        hipStream_t s);
// sum over cells and variables of (R_v / lambda / s_v)^2 into out[0] (double, deterministic)
void launch_explicit_norm(int nc, const double* res, const double* lam, Ref ref, double* partial, double* out,
                          hipStream_t s);
// features (bf16, n x kF); rnorm2 = device pointer to the sum written by
// launch_explicit_norm, rnorm2[1] = the previous step's sum; du_prev = the
// previous step's update (AoS, device order) when have_prev, else ignored.
//   0-4   p / p_inf - 1, u / U, v / U, w / U, T / T_inf - 1
//   5-9   -R / lambda / s, divided by its RMS over the mesh (rnorm)
//   10    log10(rnorm) / 8         11  freestream Mach
//   12    ln(V^(1/3) / mean)       13  log10(dt (U + a) / V^(1/3)) / 3
//   14    1                        15-19 previous update / (s rnorm)
//   20    ln(rnorm / previous rnorm)  21  1 when a previous update exists
void launch_features(int nc, const double* cs, long long ld, const double* res, const double* lam, const double* dt,
                     Ref ref, const double* rnorm2, const double* du_prev, int have_prev, bf16* x, hipStream_t s);
// The network predicts the change of the update (skip connection in FP64):
// targets t = (du - du_prev) / (s rnorm) and x0 = du_prev + y s rnorm, with
// du_prev = 0 when have_prev == 0 (du, du_prev, x0 AoS, 5 per cell).
void launch_targets(int nc, const double* du, const double* du_prev, int have_prev, Ref ref, const double* rnorm2,
                    float* t, hipStream_t s);
void launch_prediction(int nc, const float* y, const double* du_prev, int have_prev, Ref ref, const double* rnorm2,
                       double* x0, hipStream_t s);

// ---- launchers (gnn.hip) ----
// C[n x m] = A[n x k] W[m x k]^T, bf16 inputs (row-major), fp32 output; k % 32 == 0, m % 16 == 0.
void launch_gemm_nt(int n, int m, int k, const bf16* A, const bf16* W, float* C, hipStream_t s);
// Partial sums of C[m x k] = sum_rows dY[row][m] A[row][k] over row chunks of `chunk` rows:
// partial[b][m][k]; dY fp32 (converted to bf16 for the MFMA), A bf16.
int gemm_tn_blocks(int n, int chunk);
void launch_gemm_tn(int n, int m, int k, int chunk, const float* dY, const bf16* A, float* partial, hipStream_t s);
// out[i] (+)= sum_b partial[b][i], i < len, fixed order
void launch_reduce_partials(int nblocks, int len, const float* partial, float* out, int accumulate, hipStream_t s);
// edge layer forward: h_i = mean relu(P_i + Q_j + G g_e + b); writes h (bf16 or fp32) and, when Wo != null, y
void launch_edge_forward(const Graph& g, const float* pq, const float* Gm, const float* b, bf16* h_bf16, float* h_f32,
                         const float* Wo, const float* bo, float* y, hipStream_t s);
// loss + output-layer backward: dy = 2 (y - t) / (n kO); dh2 = Wo^T dy; partial grads of Wo, bo; loss partials
int out_backward_blocks(int n);
void launch_out_backward(int n, const float* y, const float* t, const float* h2, const float* Wo, float* dh2,
                         float* partial, hipStream_t s);
// edge layer backward: dPQ (n x kPQ) and per-block partials of dG (kH x kG) and db (kH)
int edge_backward_blocks(int n);
void launch_edge_backward(const Graph& g, const float* pq, const float* Gm, const float* b, const float* dh,
                          float* dpq, float* partial, hipStream_t s);
// fp32 -> bf16 copies; transpose of the layer-2 weights for dh1 = dPQ2 W2
void launch_to_bf16(int n, const float* src, bf16* dst, hipStream_t s);
void launch_transpose_to_bf16(int rows, int cols, const float* src, bf16* dst, hipStream_t s);
// Adam step over the flat parameters (bias-corrected, step t >= 1); grad is not modified
void launch_adam(int n, float* p, const float* grad, float* m, float* v, float lr, float beta1, float beta2, float eps,
                 int t, hipStream_t s);

}  // namespace gnn
}  // namespace cfdg
