// C ABI of the learned warm start (include/cfdg.h, cfdg_gnn_*): the network
// object (FP32 master parameters, BF16 copies for the MFMA products, Adam
// state), the recorded training samples with their graphs, training, and the
// hooks that the implicit step calls for features and initial guesses.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "cfdg.h"
#include "ctx.hpp"
#include "gnn.hpp"

namespace gn = cfdg::gnn;
using cfdg::bf16;
using cfdg::CfdgError;
using cfdg::DBuf;
using cfdg::hip_check;

struct cfdg_gnn {
    std::string err;
    hipStream_t stream = nullptr;
    DBuf<float> params, grad, adam_m, adam_v;
    DBuf<bf16> w1b, w2b, w2tb;  // [WP1; WQ1] (kPQ x kF), [WP2; WQ2] (kPQ x kH), its transpose (kH x kPQ)
    int adam_t = 0;
    bool bf16_fresh = false;

    struct GraphData {
        std::uint64_t owner = 0;  // context uid
        int n = 0;
        DBuf<int> adj_off, adj, rev;
        DBuf<float> eg;
        gn::Graph view() const { return gn::Graph{n, adj_off.get(), adj.get(), rev.get(), eg.get()}; }
    };
    struct Sample {
        int graph = 0;
        DBuf<bf16> x;
        DBuf<float> t;
    };
    std::vector<std::unique_ptr<GraphData>> graphs;
    std::vector<std::unique_ptr<Sample>> samples;

    // scratch sized to the largest graph used so far
    int cap = 0;
    DBuf<float> pq1, pq2, h2, y, dh2, dpq, dh1;
    DBuf<bf16> h1, dpqb;
    DBuf<float> partial;
    DBuf<double> loss_host_dev;
    float* h_partial = nullptr;

    void sync() { hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize"); }
    void check_launch() { hip_check(hipGetLastError(), "kernel launch"); }
    ~cfdg_gnn() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

std::uint64_t next_uid() {
    static std::uint64_t counter = 0;
    return ++counter;
}

template <typename F>
int guarded_g(cfdg_gnn* g, F&& fn) {
    if (!g) return CFDG_ERR_ARG;
    try {
        fn();
        return CFDG_OK;
    } catch (const CfdgError& e) {
        g->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g->err = e.what();
        return CFDG_ERR_ARG;
    }
}

template <typename F>
int guarded_c(cfdg_ctx* c, F&& fn) {
    if (!c) return CFDG_ERR_ARG;
    try {
        fn();
        return CFDG_OK;
    } catch (const CfdgError& e) {
        c->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        c->err = e.what();
        return CFDG_ERR_ARG;
    }
}

std::uint64_t splitmix64(std::uint64_t& s) {
    std::uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

double uniform01(std::uint64_t& s) { return static_cast<double>(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

// Glorot-uniform weights (bound sqrt(6 / (fan_in + fan_out))), zero biases;
// the unused input columns of W1 are zero.
std::vector<float> init_params(std::uint64_t seed) {
    std::vector<float> p(gn::kNumParams, 0.f);
    std::uint64_t s = seed;
    const auto fill = [&](int off, int rows, int cols, int fan_in, int used_cols) {
        const double bound = std::sqrt(6.0 / (fan_in + rows));
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c)
                p[static_cast<std::size_t>(off + r * cols + c)] =
                    (c < used_cols) ? static_cast<float>((2.0 * uniform01(s) - 1.0) * bound) : 0.f;
    };
    fill(gn::kOffW1, gn::kPQ, gn::kF, 2 * gn::kFUsed + gn::kG, gn::kFUsed);
    fill(gn::kOffG1, gn::kH, gn::kG, 2 * gn::kFUsed + gn::kG, gn::kG);
    fill(gn::kOffW2, gn::kPQ, gn::kH, 2 * gn::kH + gn::kG, gn::kH);
    fill(gn::kOffG2, gn::kH, gn::kG, 2 * gn::kH + gn::kG, gn::kG);
    fill(gn::kOffWo, gn::kO, gn::kH, gn::kH, gn::kH);
    return p;
}

void refresh_bf16(cfdg_gnn& g) {
    if (g.bf16_fresh) return;
    const float* p = g.params.get();
    gn::launch_to_bf16(gn::kPQ * gn::kF, p + gn::kOffW1, g.w1b.get(), g.stream);
    gn::launch_to_bf16(gn::kPQ * gn::kH, p + gn::kOffW2, g.w2b.get(), g.stream);
    gn::launch_transpose_to_bf16(gn::kPQ, gn::kH, p + gn::kOffW2, g.w2tb.get(), g.stream);
    g.bf16_fresh = true;
}

void ensure_capacity(cfdg_gnn& g, int n) {
    if (n <= g.cap) return;
    const std::size_t N = static_cast<std::size_t>(n);
    g.pq1.alloc(N * gn::kPQ);
    g.pq2.alloc(N * gn::kPQ);
    g.h2.alloc(N * gn::kH);
    g.y.alloc(N * gn::kO);
    g.dh2.alloc(N * gn::kH);
    g.dpq.alloc(N * gn::kPQ);
    g.dh1.alloc(N * gn::kH);
    g.h1.alloc(N * gn::kH);
    g.dpqb.alloc(N * gn::kPQ);
    g.cap = n;
}

constexpr int kTnChunk = 256;

std::size_t partial_floats(int n) {
    const std::size_t tn = static_cast<std::size_t>(gn::gemm_tn_blocks(n, kTnChunk)) * gn::kPQ * gn::kH;
    const std::size_t ob = static_cast<std::size_t>(gn::out_backward_blocks(n)) * (gn::kO * gn::kH + gn::kO + 1);
    const std::size_t eb = static_cast<std::size_t>(gn::edge_backward_blocks(n)) * (gn::kH * gn::kG + gn::kH);
    return std::max(tn, std::max(ob, eb));
}

// Forward pass; y in g.y.
void forward(cfdg_gnn& g, const gn::Graph& gr, const bf16* x) {
    ensure_capacity(g, gr.n);
    refresh_bf16(g);
    const float* p = g.params.get();
    gn::launch_gemm_nt(gr.n, gn::kPQ, gn::kF, x, g.w1b.get(), g.pq1.get(), g.stream);
    gn::launch_edge_forward(gr, g.pq1.get(), p + gn::kOffG1, p + gn::kOffB1, g.h1.get(), nullptr, nullptr, nullptr,
                            nullptr, g.stream);
    gn::launch_gemm_nt(gr.n, gn::kPQ, gn::kH, g.h1.get(), g.w2b.get(), g.pq2.get(), g.stream);
    gn::launch_edge_forward(gr, g.pq2.get(), p + gn::kOffG2, p + gn::kOffB2, nullptr, g.h2.get(), p + gn::kOffWo,
                            p + gn::kOffBo, g.y.get(), g.stream);
}

// Forward + backward of the mean squared error against t; gradient into
// g.grad (overwritten), returns the loss.
double loss_and_grad(cfdg_gnn& g, const gn::Graph& gr, const bf16* x, const float* t) {
    const int n = gr.n;
    forward(g, gr, x);
    const std::size_t need = partial_floats(n);
    if (g.partial.size() < need) g.partial.alloc(need);
    float* gp = g.grad.get();
    const float* p = g.params.get();
    // output layer + loss
    const int ob = gn::out_backward_blocks(n);
    const int obl = gn::kO * gn::kH + gn::kO + 1;
    gn::launch_out_backward(n, g.y.get(), t, g.h2.get(), p + gn::kOffWo, g.dh2.get(), g.partial.get(), g.stream);
    // dWo, dbo, loss are contiguous in the partial record: reduce into a staging copy
    DBuf<float> red;
    red.alloc(static_cast<std::size_t>(obl));
    gn::launch_reduce_partials(ob, obl, g.partial.get(), red.get(), 0, g.stream);
    hip_check(hipMemcpyAsync(gp + gn::kOffWo, red.get(), sizeof(float) * (gn::kO * gn::kH + gn::kO),
                             hipMemcpyDeviceToDevice, g.stream),
              "copy");
    float loss = 0.f;
    hip_check(hipMemcpyAsync(&loss, red.get() + gn::kO * gn::kH + gn::kO, sizeof(float), hipMemcpyDeviceToHost,
                             g.stream),
              "copy");
    // layer 2
    const int eb = gn::edge_backward_blocks(n);
    const int ebl = gn::kH * gn::kG + gn::kH;
    gn::launch_edge_backward(gr, g.pq2.get(), p + gn::kOffG2, p + gn::kOffB2, g.dh2.get(), g.dpq.get(), g.partial.get(),
                             g.stream);
    DBuf<float> red2;
    red2.alloc(static_cast<std::size_t>(ebl));
    gn::launch_reduce_partials(eb, ebl, g.partial.get(), red2.get(), 0, g.stream);
    hip_check(hipMemcpyAsync(gp + gn::kOffG2, red2.get(), sizeof(float) * ebl, hipMemcpyDeviceToDevice, g.stream),
              "copy");  // G2 then b2 are contiguous in both layouts
    const int tb = gn::gemm_tn_blocks(n, kTnChunk);
    gn::launch_gemm_tn(n, gn::kPQ, gn::kH, kTnChunk, g.dpq.get(), g.h1.get(), g.partial.get(), g.stream);
    gn::launch_reduce_partials(tb, gn::kPQ * gn::kH, g.partial.get(), gp + gn::kOffW2, 0, g.stream);
    // dh1 = dPQ2 [WP2; WQ2]
    gn::launch_to_bf16(n * gn::kPQ, g.dpq.get(), g.dpqb.get(), g.stream);
    gn::launch_gemm_nt(n, gn::kH, gn::kPQ, g.dpqb.get(), g.w2tb.get(), g.dh1.get(), g.stream);
    // layer 1
    gn::launch_edge_backward(gr, g.pq1.get(), p + gn::kOffG1, p + gn::kOffB1, g.dh1.get(), g.dpq.get(), g.partial.get(),
                             g.stream);
    gn::launch_reduce_partials(eb, ebl, g.partial.get(), red2.get(), 0, g.stream);
    hip_check(hipMemcpyAsync(gp + gn::kOffG1, red2.get(), sizeof(float) * ebl, hipMemcpyDeviceToDevice, g.stream),
              "copy");
    gn::launch_gemm_tn(n, gn::kPQ, gn::kF, kTnChunk, g.dpq.get(), x, g.partial.get(), g.stream);
    gn::launch_reduce_partials(tb, gn::kPQ * gn::kF, g.partial.get(), gp + gn::kOffW1, 0, g.stream);
    g.check_launch();
    g.sync();
    // the unused input columns of W1 carry zero features: their gradient is zero already
    return static_cast<double>(loss);
}

cfdg::gnn::Ref make_ref(const cfdg_ctx& c) {
    if (!c.have_ref) throw CfdgError(CFDG_ERR_STATE, "learned warm start: no freestream (cfdg_init_uniform or cfdg_set_freestream)");
    gn::Ref r{};
    const double g = c.num.gamma, R = c.num.gas_constant;
    r.rho = c.ref_rho;
    r.p = c.ref_p;
    r.T = c.ref_p / (c.ref_rho * R);
    r.a = std::sqrt(g * c.ref_p / c.ref_rho);
    const double sp = std::sqrt(c.ref_vel[0] * c.ref_vel[0] + c.ref_vel[1] * c.ref_vel[1] + c.ref_vel[2] * c.ref_vel[2]);
    r.mach = sp / r.a;
    r.speed = sp > 1.0e-12 * r.a ? sp : r.a;
    r.s[0] = r.rho;
    r.s[1] = r.s[2] = r.s[3] = r.rho * r.speed;
    r.s[4] = r.p / (g - 1.0) + 0.5 * r.rho * r.speed * r.speed;
    r.h_mean = c.h_mean;
    return r;
}

void ensure_ctx_graph(cfdg_ctx& c) {
    if (c.have_graph) return;
    const std::size_t ne = c.cf.size();
    c.gadj.alloc(ne);
    c.grev.alloc(ne);
    c.geg.alloc(ne * gn::kG);
    gn::launch_build_graph(c.nc, c.ninner, c.cf_off.get(), c.cf.get(), c.fp.get(), c.fld, c.fcells.get(), c.cs.get(),
                           c.ld, c.gadj.get(), c.grev.get(), c.geg.get(), c.stream);
    c.check_launch();
    // mean V^(1/3) (host, from the static volumes)
    const std::vector<double> planes = c.download_planes();
    double s = 0.0;
    for (int i = 0; i < c.nc; ++i) s += std::cbrt(planes[static_cast<std::size_t>(cfdg::plane_at(cfdg::kRVol, c.ld, i))]);
    c.h_mean = c.nc > 0 ? s / c.nc : 1.0;
    c.gx.alloc(static_cast<std::size_t>(c.nc) * gn::kF);
    c.gnorm.alloc(2 + 512);
    c.gnorm.zero(c.stream);
    c.du_prev.alloc(static_cast<std::size_t>(c.nc) * 5);
    c.sync();
    c.have_graph = true;
}

gn::Graph ctx_graph(const cfdg_ctx& c) {
    return gn::Graph{c.nc, c.cf_off.get(), c.gadj.get(), c.grev.get(), c.geg.get()};
}

}  // namespace

namespace cfdg {

void gnn_step_features(cfdg_ctx& c) {
    c.have_features = false;
    if (!c.gnn_record && !c.warm) return;
    ensure_ctx_graph(c);
    const gn::Ref r = make_ref(c);
    // keep the previous step's norm in gnorm[1] and its update in du_prev, then
    // form this step's norm in gnorm[0]
    hip_check(hipMemcpyAsync(c.gnorm.get() + 1, c.gnorm.get(), sizeof(double), hipMemcpyDeviceToDevice, c.stream), "copy");
    c.features_had_prev = c.have_prev_du;
    if (c.have_prev_du)
        hip_check(hipMemcpyAsync(c.du_prev.get(), c.x.get(), sizeof(double) * 5 * static_cast<std::size_t>(c.nc),
                                 hipMemcpyDeviceToDevice, c.stream),
                  "copy");
    gn::launch_explicit_norm(c.nc, c.res.get(), c.lam.get(), r, c.gnorm.get() + 2, c.gnorm.get(), c.stream);
    gn::launch_features(c.nc, c.cs.get(), c.ld, c.res.get(), c.lam.get(), c.dt.get(), r, c.gnorm.get(),
                        c.du_prev.get(), c.features_had_prev ? 1 : 0, c.gx.get(), c.stream);
    c.check_launch();
    c.have_features = true;
}

bool gnn_warm_guess(cfdg_ctx& c, double* x0) {
    if (!c.warm) return false;
    if (!c.have_features) gnn_step_features(c);
    cfdg_gnn& g = *c.warm;
    c.sync();
    forward(g, ctx_graph(c), c.gx.get());
    g.sync();
    gn::launch_prediction(c.nc, g.y.get(), c.du_prev.get(), c.features_had_prev ? 1 : 0, make_ref(c), c.gnorm.get(), x0,
                          c.stream);
    c.check_launch();
    return true;
}

}  // namespace cfdg

extern "C" {

int cfdg_gnn_create(uint64_t seed, cfdg_gnn** out) {
    if (!out) return CFDG_ERR_ARG;
    *out = nullptr;
    auto* g = new cfdg_gnn();
    const int rc = guarded_g(g, [&] {
        hip_check(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking), "hipStreamCreate");
        const std::vector<float> p = init_params(seed);
        g->params.upload(p, g->stream);
        g->grad.alloc(p.size());
        g->adam_m.alloc(p.size());
        g->adam_v.alloc(p.size());
        g->grad.zero(g->stream);
        g->adam_m.zero(g->stream);
        g->adam_v.zero(g->stream);
        g->w1b.alloc(static_cast<std::size_t>(gn::kPQ) * gn::kF);
        g->w2b.alloc(static_cast<std::size_t>(gn::kPQ) * gn::kH);
        g->w2tb.alloc(static_cast<std::size_t>(gn::kPQ) * gn::kH);
        g->sync();
    });
    if (rc != CFDG_OK) {
        delete g;
        return rc;
    }
    *out = g;
    return CFDG_OK;
}

int cfdg_gnn_destroy(cfdg_gnn* g) {
    if (!g) return CFDG_ERR_ARG;
    delete g;
    return CFDG_OK;
}

const char* cfdg_gnn_last_error(const cfdg_gnn* g) { return g ? g->err.c_str() : "null network"; }

int cfdg_gnn_num_params(int32_t* n) {
    if (!n) return CFDG_ERR_ARG;
    *n = gn::kNumParams;
    return CFDG_OK;
}

int cfdg_gnn_get_params(cfdg_gnn* g, float* p, int32_t n) {
    return guarded_g(g, [&] {
        if (!p || n != gn::kNumParams) throw CfdgError(CFDG_ERR_ARG, "get_params: buffer of cfdg_gnn_num_params floats");
        hip_check(hipMemcpyAsync(p, g->params.get(), sizeof(float) * gn::kNumParams, hipMemcpyDeviceToHost, g->stream),
                  "copy");
        g->sync();
    });
}

int cfdg_gnn_set_params(cfdg_gnn* g, const float* p, int32_t n) {
    return guarded_g(g, [&] {
        if (!p || n != gn::kNumParams) throw CfdgError(CFDG_ERR_ARG, "set_params: buffer of cfdg_gnn_num_params floats");
        std::vector<float> h(p, p + gn::kNumParams);
        for (int r = 0; r < gn::kPQ; ++r)
            for (int c = gn::kFUsed; c < gn::kF; ++c) h[static_cast<std::size_t>(gn::kOffW1 + r * gn::kF + c)] = 0.f;
        hip_check(hipMemcpyAsync(g->params.get(), h.data(), sizeof(float) * gn::kNumParams, hipMemcpyHostToDevice,
                                 g->stream),
                  "copy");
        g->adam_m.zero(g->stream);
        g->adam_v.zero(g->stream);
        g->adam_t = 0;
        g->bf16_fresh = false;
        g->sync();
    });
}

int cfdg_set_freestream(cfdg_ctx* c, double rho, const double velocity[3], double p) {
    return guarded_c(c, [&] {
        if (!velocity || !(rho > 0.0) || !(p > 0.0)) throw CfdgError(CFDG_ERR_ARG, "freestream: rho > 0, p > 0, velocity");
        c->ref_rho = rho;
        c->ref_p = p;
        for (int k = 0; k < 3; ++k) c->ref_vel[k] = velocity[k];
        c->have_ref = true;
    });
}

int cfdg_gnn_set_recording(cfdg_ctx* c, int32_t on) {
    return guarded_c(c, [&] { c->gnn_record = on != 0; });
}

int cfdg_gnn_record(cfdg_gnn* g, cfdg_ctx* c) {
    if (!c) return CFDG_ERR_ARG;
    return guarded_g(g, [&] {
        if (!c->have_features) throw CfdgError(CFDG_ERR_STATE, "record: no features (enable recording, then take an implicit step)");
        if (c->uid == 0) c->uid = next_uid();
        int gi = -1;
        for (std::size_t k = 0; k < g->graphs.size(); ++k)
            if (g->graphs[k]->owner == c->uid) gi = static_cast<int>(k);
        const std::size_t ne = c->cf.size();
        const std::size_t n = static_cast<std::size_t>(c->nc);
        c->sync();
        if (gi < 0) {
            auto gd = std::make_unique<cfdg_gnn::GraphData>();
            gd->owner = c->uid;
            gd->n = c->nc;
            gd->adj_off.alloc(n + 1);
            gd->adj.alloc(ne);
            gd->rev.alloc(ne);
            gd->eg.alloc(ne * gn::kG);
            hip_check(hipMemcpy(gd->adj_off.get(), c->cf_off.get(), (n + 1) * sizeof(int), hipMemcpyDeviceToDevice), "copy");
            hip_check(hipMemcpy(gd->adj.get(), c->gadj.get(), ne * sizeof(int), hipMemcpyDeviceToDevice), "copy");
            hip_check(hipMemcpy(gd->rev.get(), c->grev.get(), ne * sizeof(int), hipMemcpyDeviceToDevice), "copy");
            hip_check(hipMemcpy(gd->eg.get(), c->geg.get(), ne * gn::kG * sizeof(float), hipMemcpyDeviceToDevice), "copy");
            g->graphs.push_back(std::move(gd));
            gi = static_cast<int>(g->graphs.size()) - 1;
        }
        auto smp = std::make_unique<cfdg_gnn::Sample>();
        smp->graph = gi;
        smp->x.alloc(n * gn::kF);
        smp->t.alloc(n * gn::kO);
        hip_check(hipMemcpy(smp->x.get(), c->gx.get(), n * gn::kF * sizeof(bf16), hipMemcpyDeviceToDevice), "copy");
        gn::launch_targets(c->nc, c->x.get(), c->du_prev.get(), c->features_had_prev ? 1 : 0, make_ref(*c),
                           c->gnorm.get(), smp->t.get(), c->stream);
        c->check_launch();
        c->sync();
        g->samples.push_back(std::move(smp));
    });
}

int cfdg_gnn_num_samples(cfdg_gnn* g, int32_t* n) {
    return guarded_g(g, [&] {
        if (!n) throw CfdgError(CFDG_ERR_ARG, "null output");
        *n = static_cast<int32_t>(g->samples.size());
    });
}

int cfdg_gnn_loss_grad(cfdg_gnn* g, int32_t sample, double* loss, float* grad, int32_t n) {
    return guarded_g(g, [&] {
        if (sample < 0 || sample >= static_cast<int32_t>(g->samples.size())) throw CfdgError(CFDG_ERR_ARG, "no such sample");
        if (grad && n != gn::kNumParams) throw CfdgError(CFDG_ERR_ARG, "gradient buffer of cfdg_gnn_num_params floats");
        const auto& s = *g->samples[static_cast<std::size_t>(sample)];
        const double L = loss_and_grad(*g, g->graphs[static_cast<std::size_t>(s.graph)]->view(), s.x.get(), s.t.get());
        if (loss) *loss = L;
        if (grad) {
            hip_check(hipMemcpy(grad, g->grad.get(), sizeof(float) * gn::kNumParams, hipMemcpyDeviceToHost), "copy");
        }
    });
}

int cfdg_gnn_train(cfdg_gnn* g, int32_t epochs, float lr, uint64_t seed, double* mean_loss) {
    return guarded_g(g, [&] {
        if (epochs < 1 || !(lr > 0.f)) throw CfdgError(CFDG_ERR_ARG, "epochs >= 1 and lr > 0 required");
        if (g->samples.empty()) throw CfdgError(CFDG_ERR_STATE, "train: no recorded samples");
        std::vector<int> order(g->samples.size());
        std::uint64_t st = seed;
        double last = 0.0;
        for (int ep = 0; ep < epochs; ++ep) {
            std::iota(order.begin(), order.end(), 0);
            for (std::size_t i = order.size(); i > 1; --i) {  // Fisher-Yates with splitmix64
                const std::size_t j = static_cast<std::size_t>(splitmix64(st) % i);
                std::swap(order[i - 1], order[j]);
            }
            double sum = 0.0;
            for (const int k : order) {
                const auto& s = *g->samples[static_cast<std::size_t>(k)];
                sum += loss_and_grad(*g, g->graphs[static_cast<std::size_t>(s.graph)]->view(), s.x.get(), s.t.get());
                ++g->adam_t;
                gn::launch_adam(gn::kNumParams, g->params.get(), g->grad.get(), g->adam_m.get(), g->adam_v.get(), lr,
                                0.9f, 0.999f, 1.0e-8f, g->adam_t, g->stream);
                g->bf16_fresh = false;
            }
            last = sum / static_cast<double>(order.size());
        }
        g->check_launch();
        g->sync();
        if (mean_loss) *mean_loss = last;
    });
}

int cfdg_gnn_get_sample(cfdg_gnn* g, int32_t sample, int32_t* n_nodes, int32_t* n_edges, float* x, float* t,
                        int32_t* adj_off, int32_t* adj, int32_t* rev, float* eg) {
    return guarded_g(g, [&] {
        if (sample < 0 || sample >= static_cast<int32_t>(g->samples.size())) throw CfdgError(CFDG_ERR_ARG, "no such sample");
        const auto& s = *g->samples[static_cast<std::size_t>(sample)];
        const auto& gd = *g->graphs[static_cast<std::size_t>(s.graph)];
        const std::size_t n = static_cast<std::size_t>(gd.n), ne = gd.adj.size();
        if (n_nodes) *n_nodes = gd.n;
        if (n_edges) *n_edges = static_cast<int32_t>(ne);
        if (x) {
            std::vector<bf16> h(n * gn::kF);
            hip_check(hipMemcpy(h.data(), s.x.get(), h.size() * sizeof(bf16), hipMemcpyDeviceToHost), "copy");
            for (std::size_t i = 0; i < h.size(); ++i) x[i] = cfdg::bf16_to_float(h[i]);
        }
        if (t) hip_check(hipMemcpy(t, s.t.get(), n * gn::kO * sizeof(float), hipMemcpyDeviceToHost), "copy");
        if (adj_off) hip_check(hipMemcpy(adj_off, gd.adj_off.get(), (n + 1) * sizeof(int), hipMemcpyDeviceToHost), "copy");
        if (adj) hip_check(hipMemcpy(adj, gd.adj.get(), ne * sizeof(int), hipMemcpyDeviceToHost), "copy");
        if (rev) hip_check(hipMemcpy(rev, gd.rev.get(), ne * sizeof(int), hipMemcpyDeviceToHost), "copy");
        if (eg) hip_check(hipMemcpy(eg, gd.eg.get(), ne * gn::kG * sizeof(float), hipMemcpyDeviceToHost), "copy");
    });
}

int cfdg_gnn_predict(cfdg_gnn* g, cfdg_ctx* c, double* du) {
    if (!c) return CFDG_ERR_ARG;
    return guarded_g(g, [&] {
        if (!c->have_dt) throw CfdgError(CFDG_ERR_STATE, "predict needs eval_residual and compute_dt");
        cfdg_gnn* keep = c->warm;
        c->warm = g;
        c->have_features = false;
        DBuf<double> x0;
        x0.alloc(static_cast<std::size_t>(c->nc) * 5);
        try {
            cfdg::gnn_warm_guess(*c, x0.get());
        } catch (...) {
            c->warm = keep;
            throw;
        }
        c->warm = keep;
        std::vector<double> h(static_cast<std::size_t>(c->nc) * 5);
        c->d2h(h.data(), x0.get(), h.size());
        c->sync();
        if (du) c->vec_to_mesh(h, du);
    });
}

int cfdg_set_warm_start(cfdg_ctx* c, cfdg_gnn* g) {
    return guarded_c(c, [&] {
        c->warm = g;
        c->warm_used = c->warm_rejected = 0;
    });
}

int cfdg_warm_start_stats(cfdg_ctx* c, int64_t* used, int64_t* rejected) {
    return guarded_c(c, [&] {
        if (used) *used = c->warm_used;
        if (rejected) *rejected = c->warm_rejected;
    });
}

}  // extern "C"
