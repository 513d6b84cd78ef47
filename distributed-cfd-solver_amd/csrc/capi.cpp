// C ABI of the library (include/cfdg.h): context set-up, device memory and
// the orchestration of the residual, Jacobian and BiCGSTAB stages on one HIP
// stream.  Every entry point catches exceptions and turns them into status codes.
#include "cfdg.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "kernels_linalg.hpp"
#include "mesh.hpp"
#include "setup.hpp"

struct cfdg_mesh {
    cfdg::Mesh m;
};

namespace {

struct CfdgError : std::runtime_error {
    int code;
    CfdgError(int c, const std::string& w) : std::runtime_error(w), code(c) {}
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw CfdgError(CFDG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
class DBuf {
  public:
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void alloc(std::size_t n) {
        release();
        n_ = n;
        if (n > 0) hip_check(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)), "hipMalloc");
    }
    void upload(const std::vector<T>& h, hipStream_t s) {
        alloc(h.size());
        if (n_ > 0) hip_check(hipMemcpyAsync(p_, h.data(), n_ * sizeof(T), hipMemcpyHostToDevice, s), "upload");
    }
    void zero(hipStream_t s) {
        if (n_ > 0) hip_check(hipMemsetAsync(p_, 0, n_ * sizeof(T), s), "memset");
    }
    T* get() const { return p_; }
    std::size_t size() const { return n_; }

  private:
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* p_ = nullptr;
    std::size_t n_ = 0;
};

// Resolved boundary parameters (bc_manager.hpp:235-352 builders + resolve_inflow :37-111).
cfdg::BcDev resolve_bc(const cfdg_bc& d, double gamma, double R) {
    cfdg::BcDev b{};
    b.type = d.type;
    const auto inflow = [&](double p_ref, double& u, double& v, double& w) {
        if (d.inflow_mode == CFDG_INFLOW_VELOCITY) {
            u = d.velocity[0];
            v = d.velocity[1];
            w = d.velocity[2];
            return;
        }
        const double rho = p_ref / (R * d.t);
        const double a = std::sqrt(gamma * p_ref / rho);
        const double vm = d.mach * a;
        if (d.inflow_mode == CFDG_INFLOW_MACH_ANGLES) {
            const double k = M_PI / 180.0;
            const double ar = d.alpha_deg * k, br = d.beta_deg * k;
            u = vm * std::cos(ar) * std::cos(br);
            v = -vm * std::sin(br);
            w = vm * std::sin(ar) * std::cos(br);
        } else if (d.inflow_mode == CFDG_INFLOW_MACH_DIRECTION) {
            const double nrm = std::sqrt(d.direction[0] * d.direction[0] + d.direction[1] * d.direction[1] +
                                         d.direction[2] * d.direction[2]);
            const double inv = (nrm > 1.0e-14) ? (1.0 / nrm) : 0.0;
            u = vm * d.direction[0] * inv;
            v = vm * d.direction[1] * inv;
            w = vm * d.direction[2] * inv;
        } else {
            throw CfdgError(CFDG_ERR_CONFIG, "unknown inflow mode");
        }
    };
    switch (d.type) {
        case CFDG_BC_SUPERSONIC_INLET:
        case CFDG_BC_FARFIELD:
            b.p = d.p;
            b.T = d.t;
            inflow(d.p, b.u, b.v, b.w);
            break;
        case CFDG_BC_NO_SLIP_WALL:
            b.u = d.velocity[0];
            b.v = d.velocity[1];
            b.w = d.velocity[2];
            b.T = d.t;
            break;
        case CFDG_BC_NO_SLIP_WALL_HEAT_FLUX:
            b.u = d.velocity[0];
            b.v = d.velocity[1];
            b.w = d.velocity[2];
            b.tmp_grad = d.tmp_grad;
            break;
        case CFDG_BC_SUBSONIC_INLET: {
            const double p_ref = (d.p > 0.0) ? d.p : 101325.0;
            b.T = d.t;
            inflow(p_ref, b.u, b.v, b.w);
            break;
        }
        case CFDG_BC_SUBSONIC_OUTLET:
            b.p = d.p;
            break;
        case CFDG_BC_SUPERSONIC_OUTLET:
        case CFDG_BC_SLIP_WALL:
        case CFDG_BC_SYMMETRY:
            break;
        default:
            throw CfdgError(CFDG_ERR_CONFIG, "unknown boundary condition type " + std::to_string(d.type));
    }
    return b;
}

}  // namespace

struct cfdg_ctx {
    std::string err;
    hipStream_t stream = nullptr;
    int nc = 0, nf = 0, ninner = 0, nbf = 0, ntot = 0, npatch = 0;
    cfdg_numerics num{};
    cfdg::NumArgs nargs{};
    cfdg::BcTable bcs{};

    // mesh
    DBuf<int> owner, neigh, cf_off, cf, ccf_off, ccf, ccn_off, ccn, bface_bc;
    DBuf<double> fnx, fny, fnz, farea, fcx, fcy, fcz, dvx, dvy, dvz, dinv, ccx, ccy, ccz, vol, lsq;
    std::int64_t n_edges = 0;
    // fields
    DBuf<double> U, q, grad, phi, res, lam, dt, alpha;
    bool have_residual = false;
    bool have_dt = false;
    // BSR
    std::vector<std::int32_t> h_row_ptr, h_cols, h_diag;
    DBuf<int> row_ptr, cols, diag;
    DBuf<double> values, lu;
    DBuf<std::uint8_t> piv;
    DBuf<int> singular;
    // SGS row groups: 0 multicolour fwd, 1 multicolour bwd, 2 sequential fwd, 3 sequential bwd
    std::vector<std::int32_t> h_gofs[4];
    DBuf<int> grows[4];
    cfdg::Csr ccf_host;
    bool have_seq_groups = false;
    // Krylov work vectors (AoS, 5 nc)
    DBuf<double> b, x, r, r0, p, v, s, t, ph, sh;
    DBuf<double> partial, dots;
    double* h_dots = nullptr;  // pinned
    // timing
    hipEvent_t ev[8] = {};
    double stage_ms[6] = {0, 0, 0, 0, 0, 0};

    cfdg::MeshArgs margs() const {
        cfdg::MeshArgs a{};
        a.nc = nc;
        a.nf = nf;
        a.ninner = ninner;
        a.ld = ntot;
        a.owner = owner.get();
        a.neigh = neigh.get();
        a.fnx = fnx.get();
        a.fny = fny.get();
        a.fnz = fnz.get();
        a.farea = farea.get();
        a.fcx = fcx.get();
        a.fcy = fcy.get();
        a.fcz = fcz.get();
        a.dvx = dvx.get();
        a.dvy = dvy.get();
        a.dvz = dvz.get();
        a.dinv = dinv.get();
        a.ccx = ccx.get();
        a.ccy = ccy.get();
        a.ccz = ccz.get();
        a.vol = vol.get();
        a.cf_off = cf_off.get();
        a.cf = cf.get();
        a.ccf_off = ccf_off.get();
        a.ccf = ccf.get();
        a.ccn_off = ccn_off.get();
        a.ccn = ccn.get();
        a.lsq = lsq.get();
        a.n_edges = n_edges;
        a.bface_bc = bface_bc.get();
        return a;
    }
    cfdg::FieldArgs fargs() const {
        cfdg::FieldArgs f{};
        f.U = U.get();
        f.q = q.get();
        f.grad = grad.get();
        f.phi = phi.get();
        f.res = res.get();
        f.lam = lam.get();
        return f;
    }
    cfdg::BsrArgs bargs() const {
        cfdg::BsrArgs a{};
        a.n = nc;
        a.row_ptr = row_ptr.get();
        a.cols = cols.get();
        a.diag = diag.get();
        return a;
    }
    void sync() { hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize"); }
    void check_launch() { hip_check(hipGetLastError(), "kernel launch"); }

    ~cfdg_ctx() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (h_dots) (void)hipHostFree(h_dots);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <typename F>
int guarded(cfdg_ctx* c, F&& fn) {
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

void build_seq_groups(cfdg_ctx& c) {
    if (c.have_seq_groups) return;
    std::int32_t nf = 0, nb = 0;
    const auto lf = cfdg::forward_levels(c.ccf_host, c.nc, nf);
    const auto lb = cfdg::backward_levels(c.ccf_host, c.nc, nb);
    const auto gf = cfdg::groups_from_labels(lf, nf, false);
    const auto gb = cfdg::groups_from_labels(lb, nb, false);
    c.h_gofs[2] = gf.off;
    c.h_gofs[3] = gb.off;
    c.grows[2].upload(gf.rows, c.stream);
    c.grows[3].upload(gb.rows, c.stream);
    c.sync();
    c.have_seq_groups = true;
}

void eval_residual(cfdg_ctx& c) {
    const auto m = c.margs();
    const auto f = c.fargs();
    hip_check(hipEventRecord(c.ev[0], c.stream), "event");
    cfdg::launch_primitives(m, f, c.nargs, c.stream);
    cfdg::launch_ghosts(m, f, c.nargs, c.bcs, c.stream);
    hip_check(hipEventRecord(c.ev[1], c.stream), "event");
    const bool need_grad = c.nargs.muscl || c.nargs.viscous;
    if (need_grad) {
        cfdg::launch_gradient_limiter(m, f, c.nargs, c.stream);
        hip_check(hipEventRecord(c.ev[2], c.stream), "event");
        cfdg::launch_ghost_gradients(m, f, c.nargs, c.bcs, c.stream);
    } else {
        hip_check(hipEventRecord(c.ev[2], c.stream), "event");
    }
    hip_check(hipEventRecord(c.ev[3], c.stream), "event");
    cfdg::launch_flux(m, f, c.nargs, c.stream);
    hip_check(hipEventRecord(c.ev[4], c.stream), "event");
    c.check_launch();
    c.sync();
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, c.ev[0], c.ev[1]), "elapsed");
    c.stage_ms[0] = ms;
    hip_check(hipEventElapsedTime(&ms, c.ev[1], c.ev[2]), "elapsed");
    c.stage_ms[1] = ms;
    hip_check(hipEventElapsedTime(&ms, c.ev[2], c.ev[3]), "elapsed");
    c.stage_ms[2] = ms;
    hip_check(hipEventElapsedTime(&ms, c.ev[3], c.ev[4]), "elapsed");
    c.stage_ms[3] = ms;
    c.have_residual = true;
}

void compute_dt(cfdg_ctx& c) {
    if (!c.have_residual) throw CfdgError(CFDG_ERR_STATE, "compute_dt before eval_residual");
    cfdg::launch_dt(c.margs(), c.lam.get(), c.num.cfl, c.dt.get(), c.alpha.get(), c.stream);
    c.check_launch();
    c.sync();
    c.have_dt = true;
}

void assemble_jacobian(cfdg_ctx& c) {
    if (!c.have_dt) throw CfdgError(CFDG_ERR_STATE, "assemble_jacobian before compute_dt");
    hip_check(hipEventRecord(c.ev[5], c.stream), "event");
    cfdg::launch_jacobian(c.margs(), c.fargs(), c.nargs, c.bargs(), c.alpha.get(), c.values.get(), c.stream);
    hip_check(hipEventRecord(c.ev[6], c.stream), "event");
    c.check_launch();
    c.sync();
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, c.ev[5], c.ev[6]), "elapsed");
    c.stage_ms[4] = ms;
}

void fetch_dots(cfdg_ctx& c, int n) {
    hip_check(hipMemcpyAsync(c.h_dots, c.dots.get(), static_cast<std::size_t>(n) * sizeof(double), hipMemcpyDeviceToHost,
                             c.stream),
              "dots D2H");
    c.sync();
}

double dot_device(cfdg_ctx& c, const double* a, const double* b) {
    cfdg::DotPairs pr{};
    pr.a[0] = a;
    pr.b[0] = b;
    cfdg::launch_dots(5LL * c.nc, 1, pr, c.partial.get(), c.dots.get(), c.stream);
    fetch_dots(c, 1);
    return c.h_dots[0];
}

struct Precond {
    int kind = CFDG_PRECOND_NONE;
    int order = CFDG_SGS_MULTICOLOUR;
};

// z = M^-1 r: z = 0, forward groups, backward groups (preconditioners.cpp:171-194)
void precond_apply(cfdg_ctx& c, const Precond& P, const double* r, double* z) {
    const long long n5 = 5LL * c.nc;
    if (P.kind == CFDG_PRECOND_NONE) {
        hip_check(hipMemcpyAsync(z, r, static_cast<std::size_t>(n5) * sizeof(double), hipMemcpyDeviceToDevice, c.stream),
                  "copy");
        return;
    }
    hip_check(hipMemsetAsync(z, 0, static_cast<std::size_t>(n5) * sizeof(double), c.stream), "memset");
    const int gf = (P.order == CFDG_SGS_SEQUENTIAL) ? 2 : 0;
    const auto b = c.bargs();
    for (int sweep = 0; sweep < 2; ++sweep) {
        const int gi = gf + sweep;
        const auto& off = c.h_gofs[gi];
        for (std::size_t k = 0; k + 1 < off.size(); ++k) {
            cfdg::launch_sgs_group(c.grows[gi].get() + off[k], off[k + 1] - off[k], b, c.values.get(), c.lu.get(),
                                   c.piv.get(), r, z, c.stream);
        }
    }
}

void precond_setup(cfdg_ctx& c, const Precond& P) {
    if (P.kind != CFDG_PRECOND_SGS) return;
    if (P.order == CFDG_SGS_SEQUENTIAL) build_seq_groups(c);
    c.singular.zero(c.stream);
    cfdg::launch_lu_factor(c.nc, c.diag.get(), c.values.get(), c.lu.get(), c.piv.get(), c.singular.get(), c.stream);
    c.check_launch();
    int sing = 0;
    hip_check(hipMemcpyAsync(&sing, c.singular.get(), sizeof(int), hipMemcpyDeviceToHost, c.stream), "D2H");
    c.sync();
    if (sing != 0) throw CfdgError(CFDG_ERR_SINGULAR, "SGS: singular diagonal block at row " + std::to_string(sing - 1));
}

// Right-preconditioned BiCGSTAB (bicgstab.cpp:42-201) on device vectors; x holds
// the initial guess on entry and the solution on exit.
cfdg_iter_result bicgstab(cfdg_ctx& c, const cfdg_linparams& lp, const Precond& P, double* x, const double* bvec) {
    const long long N = 5LL * c.nc;
    const std::size_t bytes = static_cast<std::size_t>(N) * sizeof(double);
    cfdg_iter_result res{CFDG_MAX_ITERATIONS, 0, 0.0, 0.0};
    const auto spmv = [&](const double* in, double* out) { cfdg::launch_spmv(c.bargs(), c.values.get(), in, out, c.stream); };

    spmv(x, c.v.get());
    cfdg::launch_axpy_copy(N, bvec, -1.0, c.v.get(), c.r.get(), c.stream);
    {
        cfdg::DotPairs pr{};
        pr.a[0] = bvec;
        pr.b[0] = bvec;
        pr.a[1] = c.r.get();
        pr.b[1] = c.r.get();
        cfdg::launch_dots(N, 2, pr, c.partial.get(), c.dots.get(), c.stream);
        fetch_dots(c, 2);
    }
    const double bnorm = std::sqrt(c.h_dots[0]);
    double rnorm = std::sqrt(c.h_dots[1]);
    const double scale = (bnorm > 0.0) ? bnorm : 1.0;
    const double tol = std::max(lp.abs_tol, lp.rel_tol * scale);
    res.initial_residual = rnorm;
    if (!std::isfinite(rnorm)) {
        res.status = CFDG_DIVERGED;
    } else if (rnorm <= tol) {
        res.status = CFDG_CONVERGED;
    }
    int it = 0;
    if (res.status == CFDG_MAX_ITERATIONS) {
        hip_check(hipMemcpyAsync(c.r0.get(), c.r.get(), bytes, hipMemcpyDeviceToDevice, c.stream), "copy");
        hip_check(hipMemsetAsync(c.v.get(), 0, bytes, c.stream), "memset");
        double rho = rnorm * rnorm, rho_old = 1.0, alpha = 1.0, omega = 1.0;
        while (true) {
            if (it >= lp.max_iter) {
                res.status = CFDG_MAX_ITERATIONS;
                break;
            }
            ++it;
            if (it == 1) {
                hip_check(hipMemcpyAsync(c.p.get(), c.r.get(), bytes, hipMemcpyDeviceToDevice, c.stream), "copy");
            } else {
                const double beta = (rho / rho_old) * (alpha / omega);
                if (!std::isfinite(beta)) {
                    res.status = CFDG_DIVERGED;
                    break;
                }
                cfdg::launch_update_p(N, beta, -beta * omega, c.r.get(), c.v.get(), c.p.get(), c.stream);
            }
            precond_apply(c, P, c.p.get(), c.ph.get());
            spmv(c.ph.get(), c.v.get());
            const double gamma = dot_device(c, c.r0.get(), c.v.get());
            if (!std::isfinite(gamma) || gamma == 0.0) {
                res.status = CFDG_DIVERGED;
                break;
            }
            alpha = rho / gamma;
            cfdg::launch_axpy_copy(N, c.r.get(), -alpha, c.v.get(), c.s.get(), c.stream);
            precond_apply(c, P, c.s.get(), c.sh.get());
            spmv(c.sh.get(), c.t.get());
            {
                cfdg::DotPairs pr{};
                const double* A[5] = {c.t.get(), c.t.get(), c.r0.get(), c.r0.get(), c.s.get()};
                const double* B[5] = {c.t.get(), c.s.get(), c.t.get(), c.s.get(), c.s.get()};
                for (int j = 0; j < 5; ++j) {
                    pr.a[j] = A[j];
                    pr.b[j] = B[j];
                }
                cfdg::launch_dots(N, 5, pr, c.partial.get(), c.dots.get(), c.stream);
                fetch_dots(c, 5);
            }
            const double tt = c.h_dots[0], ts = c.h_dots[1], r0t = c.h_dots[2], r0s = c.h_dots[3], ss = c.h_dots[4];
            if (!std::isfinite(ts) || !std::isfinite(tt) || tt == 0.0) {
                res.status = CFDG_DIVERGED;
                break;
            }
            const double omega_new = ts / tt;
            cfdg::launch_update_xr(N, alpha, omega_new, c.ph.get(), c.sh.get(), c.s.get(), c.t.get(), x, c.r.get(),
                                   c.stream);
            const double rn2 = ss - 2.0 * omega_new * ts + omega_new * omega_new * tt;
            rnorm = std::sqrt(std::max(rn2, 0.0));
            const double rho_new = r0s - omega_new * r0t;
            if (!std::isfinite(rnorm) || !std::isfinite(omega_new) || omega_new == 0.0) {
                res.status = CFDG_DIVERGED;
                break;
            }
            if (it % 32 == 0) rnorm = std::sqrt(dot_device(c, c.r.get(), c.r.get()));
            if (rnorm <= tol) {
                res.status = CFDG_CONVERGED;
                break;
            }
            if (!std::isfinite(rho_new) || rho_new == 0.0) {
                res.status = CFDG_DIVERGED;
                break;
            }
            rho_old = rho;
            rho = rho_new;
            omega = omega_new;
        }
    }
    c.check_launch();
    c.sync();
    res.iterations = it;
    res.final_residual = rnorm;
    return res;
}

void linear_solve(cfdg_ctx& c, const cfdg_linparams& lp, const double* hb, const double* hx0, double* hx,
                  cfdg_iter_result* out) {
    if (lp.preconditioner != CFDG_PRECOND_NONE && lp.preconditioner != CFDG_PRECOND_SGS)
        throw CfdgError(CFDG_ERR_CONFIG, "unknown preconditioner");
    if (lp.sgs_order != CFDG_SGS_MULTICOLOUR && lp.sgs_order != CFDG_SGS_SEQUENTIAL)
        throw CfdgError(CFDG_ERR_CONFIG, "unknown SGS order");
    if (lp.max_iter < 0) throw CfdgError(CFDG_ERR_ARG, "max_iter must be >= 0");
    const std::size_t bytes = 5ULL * static_cast<std::size_t>(c.nc) * sizeof(double);
    if (hb) {
        hip_check(hipMemcpyAsync(c.b.get(), hb, bytes, hipMemcpyHostToDevice, c.stream), "b H2D");
    } else {
        if (!c.have_residual) throw CfdgError(CFDG_ERR_STATE, "implicit right-hand side needs a residual");
        cfdg::launch_rhs(c.nc, c.ntot, c.res.get(), c.b.get(), c.stream);
    }
    if (hx0) {
        hip_check(hipMemcpyAsync(c.x.get(), hx0, bytes, hipMemcpyHostToDevice, c.stream), "x0 H2D");
    } else {
        c.x.zero(c.stream);
    }
    Precond P;
    P.kind = lp.preconditioner;
    P.order = lp.sgs_order;
    hip_check(hipEventRecord(c.ev[5], c.stream), "event");
    precond_setup(c, P);
    const cfdg_iter_result r = bicgstab(c, lp, P, c.x.get(), c.b.get());
    hip_check(hipEventRecord(c.ev[6], c.stream), "event");
    if (hx) hip_check(hipMemcpyAsync(hx, c.x.get(), bytes, hipMemcpyDeviceToHost, c.stream), "x D2H");
    c.sync();
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, c.ev[5], c.ev[6]), "elapsed");
    c.stage_ms[5] = ms;
    if (out) *out = r;
}

void residual_norms(cfdg_ctx& c, double l2[5]) {
    if (!c.have_residual) throw CfdgError(CFDG_ERR_STATE, "residual norms before eval_residual");
    cfdg::launch_res_sq(c.nc, c.ntot, c.res.get(), c.partial.get(), c.dots.get(), c.stream);
    fetch_dots(c, 5);
    const double scale = 1.0 / static_cast<double>(std::max(c.nc, 1));
    for (int v = 0; v < 5; ++v) l2[v] = std::sqrt(c.h_dots[v] * scale);
}

cfdg::Mesh mesh_from_view(const cfdg_mesh_view& v) {
    if (v.n_cells <= 0 || v.n_faces <= 0 || v.n_inner_faces < 0 || v.n_inner_faces > v.n_faces || v.n_patches < 0)
        throw CfdgError(CFDG_ERR_ARG, "mesh view: invalid sizes");
    cfdg::Mesh m;
    const auto nc = static_cast<std::size_t>(v.n_cells), nf = static_cast<std::size_t>(v.n_faces),
               nn = static_cast<std::size_t>(v.n_nodes);
    m.n_cells = v.n_cells;
    m.n_nodes = v.n_nodes;
    m.n_faces = v.n_faces;
    m.n_inner_faces = v.n_inner_faces;
    m.n_internal_faces = v.n_inner_faces;
    m.cell_type.assign(v.cell_type, v.cell_type + nc);
    m.cell_gid.resize(nc);
    for (std::size_t i = 0; i < nc; ++i) m.cell_gid[i] = static_cast<std::int64_t>(i);
    m.cell_nodes_off.assign(v.cell_nodes_off, v.cell_nodes_off + nc + 1);
    m.cell_nodes.assign(v.cell_nodes, v.cell_nodes + m.cell_nodes_off.back());
    m.cell_cx.assign(v.cell_cx, v.cell_cx + nc);
    m.cell_cy.assign(v.cell_cy, v.cell_cy + nc);
    m.cell_cz.assign(v.cell_cz, v.cell_cz + nc);
    m.cell_vol.assign(v.cell_vol, v.cell_vol + nc);
    m.node_x.assign(v.node_x, v.node_x + nn);
    m.node_y.assign(v.node_y, v.node_y + nn);
    m.node_z.assign(v.node_z, v.node_z + nn);
    m.node_gid.resize(nn);
    for (std::size_t i = 0; i < nn; ++i) m.node_gid[i] = static_cast<std::int64_t>(i);
    m.face_owner.assign(v.face_owner, v.face_owner + nf);
    m.face_neigh.assign(v.face_neigh, v.face_neigh + nf);
    m.face_patch.assign(v.face_patch, v.face_patch + nf);
    m.face_type.assign(v.face_type, v.face_type + nf);
    m.face_nodes_off.assign(v.face_nodes_off, v.face_nodes_off + nf + 1);
    m.face_nodes.assign(v.face_nodes, v.face_nodes + m.face_nodes_off.back());
    m.face_cx.assign(v.face_cx, v.face_cx + nf);
    m.face_cy.assign(v.face_cy, v.face_cy + nf);
    m.face_cz.assign(v.face_cz, v.face_cz + nf);
    m.face_nx.assign(v.face_nx, v.face_nx + nf);
    m.face_ny.assign(v.face_ny, v.face_ny + nf);
    m.face_nz.assign(v.face_nz, v.face_nz + nf);
    m.face_area.assign(v.face_area, v.face_area + nf);
    for (std::int32_t p = 0; p < v.n_patches; ++p) m.patch_names.push_back("patch" + std::to_string(p));
    for (std::size_t i = 0; i < nc; ++i) {
        if (m.cell_nodes_off[i + 1] < m.cell_nodes_off[i]) throw CfdgError(CFDG_ERR_MESH, "cell node offsets");
    }
    for (const auto n : m.cell_nodes)
        if (n < 0 || n >= v.n_nodes) throw CfdgError(CFDG_ERR_MESH, "cell node index out of range");
    const std::string e = cfdg::check_mesh(m);
    if (!e.empty()) throw CfdgError(CFDG_ERR_MESH, e);
    return m;
}

thread_local std::string g_mesh_err;

}  // namespace

extern "C" {

int cfdg_mesh_generate(const cfdg_mesh_spec* spec, cfdg_mesh** out) {
    if (!spec || !out) return CFDG_ERR_ARG;
    *out = nullptr;
    try {
        cfdg::MeshSpec s;
        s.kind = spec->kind;
        s.ni = spec->ni;
        s.nj = spec->nj;
        s.nk = spec->nk;
        s.x0 = spec->x0;
        s.x1 = spec->x1;
        s.y0 = spec->y0;
        s.y1 = spec->y1;
        s.z0 = spec->z0;
        s.z1 = spec->z1;
        s.y_ratio = spec->y_ratio;
        s.ramp_x = spec->ramp_x;
        s.ramp_deg = spec->ramp_deg;
        s.bump_h = spec->bump_h;
        s.bump_w = spec->bump_w;
        s.bump_x = spec->bump_x;
        s.jitter = spec->jitter;
        s.seed = spec->seed;
        s.split_ymin = spec->split_ymin;
        s.split_ymin_x = spec->split_ymin_x;
        s.reorder = spec->reorder;
        auto* m = new cfdg_mesh();
        try {
            m->m = cfdg::generate_mesh(s);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
        return CFDG_OK;
    } catch (const std::exception& e) {
        g_mesh_err = e.what();
        std::fprintf(stderr, "cfdg_mesh_generate: %s\n", e.what());
        return CFDG_ERR_MESH;
    }
}

int cfdg_mesh_from_view(const cfdg_mesh_view* view, cfdg_mesh** out) {
    if (!view || !out) return CFDG_ERR_ARG;
    *out = nullptr;
    try {
        auto* m = new cfdg_mesh();
        try {
            m->m = mesh_from_view(*view);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
        return CFDG_OK;
    } catch (const CfdgError& e) {
        std::fprintf(stderr, "cfdg_mesh_from_view: %s\n", e.what());
        return e.code;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "cfdg_mesh_from_view: %s\n", e.what());
        return CFDG_ERR_MESH;
    }
}

int cfdg_mesh_get_view(const cfdg_mesh* mesh, cfdg_mesh_view* v) {
    if (!mesh || !v) return CFDG_ERR_ARG;
    const cfdg::Mesh& m = mesh->m;
    v->n_cells = m.n_cells;
    v->n_nodes = m.n_nodes;
    v->n_faces = m.n_faces;
    v->n_inner_faces = m.n_inner_faces;
    v->n_patches = static_cast<std::int32_t>(m.patch_names.size());
    v->cell_type = m.cell_type.data();
    v->cell_nodes_off = m.cell_nodes_off.data();
    v->cell_nodes = m.cell_nodes.data();
    v->cell_cx = m.cell_cx.data();
    v->cell_cy = m.cell_cy.data();
    v->cell_cz = m.cell_cz.data();
    v->cell_vol = m.cell_vol.data();
    v->node_x = m.node_x.data();
    v->node_y = m.node_y.data();
    v->node_z = m.node_z.data();
    v->face_owner = m.face_owner.data();
    v->face_neigh = m.face_neigh.data();
    v->face_patch = m.face_patch.data();
    v->face_type = m.face_type.data();
    v->face_nodes_off = m.face_nodes_off.data();
    v->face_nodes = m.face_nodes.data();
    v->face_cx = m.face_cx.data();
    v->face_cy = m.face_cy.data();
    v->face_cz = m.face_cz.data();
    v->face_nx = m.face_nx.data();
    v->face_ny = m.face_ny.data();
    v->face_nz = m.face_nz.data();
    v->face_area = m.face_area.data();
    return CFDG_OK;
}

int cfdg_mesh_patch_name(const cfdg_mesh* mesh, int32_t patch, char* buf, int32_t buflen) {
    if (!mesh || !buf || buflen <= 0) return CFDG_ERR_ARG;
    if (patch < 0 || patch >= static_cast<int32_t>(mesh->m.patch_names.size())) return CFDG_ERR_ARG;
    std::snprintf(buf, static_cast<std::size_t>(buflen), "%s", mesh->m.patch_names[static_cast<std::size_t>(patch)].c_str());
    return CFDG_OK;
}

int cfdg_mesh_reorder_rcm(cfdg_mesh* mesh) {
    if (!mesh) return CFDG_ERR_ARG;
    try {
        cfdg::apply_cell_permutation(mesh->m, cfdg::rcm_order(mesh->m));
        return CFDG_OK;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "cfdg_mesh_reorder_rcm: %s\n", e.what());
        return CFDG_ERR_MESH;
    }
}

void cfdg_mesh_free(cfdg_mesh* mesh) { delete mesh; }

int cfdg_create(const cfdg_mesh* mesh, const cfdg_numerics* num, const cfdg_bc* bcs, int32_t n_bcs, cfdg_ctx** out) {
    if (!mesh || !num || !out || (n_bcs > 0 && !bcs)) return CFDG_ERR_ARG;
    *out = nullptr;
    auto* c = new cfdg_ctx();
    const int rc = guarded(c, [&] {
        const cfdg::Mesh& m = mesh->m;
        if (n_bcs != static_cast<int32_t>(m.patch_names.size()))
            throw CfdgError(CFDG_ERR_ARG, "one boundary condition per mesh patch is required");
        if (n_bcs > cfdg::kMaxPatches) throw CfdgError(CFDG_ERR_ARG, "too many patches");
        if (num->flow_model != CFDG_INVISCID_FLOW && num->flow_model != CFDG_VISCOUS_FLOW)
            throw CfdgError(CFDG_ERR_CONFIG, "unknown flow model");
        if (num->reconstruction != CFDG_FIRST_ORDER && num->reconstruction != CFDG_MUSCL)
            throw CfdgError(CFDG_ERR_CONFIG, "unknown reconstruction");
        if (num->limiter < CFDG_LIMITER_VENKAT || num->limiter > CFDG_LIMITER_UNLIMITED)
            throw CfdgError(CFDG_ERR_CONFIG, "unknown limiter");
        if (num->gradient < CFDG_GREEN_GAUSS_FACE || num->gradient > CFDG_LEAST_SQUARES_NODE)
            throw CfdgError(CFDG_ERR_CONFIG, "unknown gradient");
        if (!(num->gamma > 1.0) || !(num->gas_constant > 0.0) || !(num->cfl > 0.0))
            throw CfdgError(CFDG_ERR_CONFIG, "gamma > 1, gas_constant > 0 and cfl > 0 required");
        if (num->flow_model == CFDG_VISCOUS_FLOW && (!(num->prandtl > 0.0) || !(num->prandtl_turb > 0.0)))
            throw CfdgError(CFDG_ERR_CONFIG, "viscous flow needs prandtl > 0 and prandtl_turb > 0");

        int ndev = 0;
        hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
        if (ndev < 1) throw CfdgError(CFDG_ERR_HIP, "no HIP device");
        hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        for (auto& e : c->ev) hip_check(hipEventCreate(&e), "hipEventCreate");
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&c->h_dots), 8 * sizeof(double)), "hipHostMalloc");

        c->num = *num;
        c->nc = m.n_cells;
        c->nf = m.n_faces;
        c->ninner = m.n_inner_faces;
        c->nbf = m.n_faces - m.n_inner_faces;
        c->ntot = c->nc + c->nbf;
        c->npatch = n_bcs;
        c->nargs.gas = cfdg::Gas{num->gamma, num->gas_constant, num->prandtl, num->prandtl_turb};
        c->nargs.viscous = num->flow_model == CFDG_VISCOUS_FLOW;
        c->nargs.muscl = num->reconstruction == CFDG_MUSCL;
        c->nargs.limiter = num->limiter;
        c->nargs.gradient = num->gradient;
        c->nargs.k3 = num->venkat_k * num->venkat_k * num->venkat_k;
        for (int32_t p = 0; p < n_bcs; ++p) c->bcs.bc[p] = resolve_bc(bcs[p], num->gamma, num->gas_constant);

        hipStream_t s = c->stream;
        c->owner.upload(m.face_owner, s);
        c->neigh.upload(m.face_neigh, s);
        c->fnx.upload(m.face_nx, s);
        c->fny.upload(m.face_ny, s);
        c->fnz.upload(m.face_nz, s);
        c->farea.upload(m.face_area, s);
        c->fcx.upload(m.face_cx, s);
        c->fcy.upload(m.face_cy, s);
        c->fcz.upload(m.face_cz, s);
        c->ccx.upload(m.cell_cx, s);
        c->ccy.upload(m.cell_cy, s);
        c->ccz.upload(m.cell_cz, s);
        c->vol.upload(m.cell_vol, s);
        {
            std::vector<double> dx, dy, dz, di;
            cfdg::face_distances(m, dx, dy, dz, di);
            c->dvx.upload(dx, s);
            c->dvy.upload(dy, s);
            c->dvz.upload(dz, s);
            c->dinv.upload(di, s);
            c->sync();
        }
        {
            const cfdg::Csr cf = cfdg::cell_faces(m);
            c->cf_off.upload(cf.off, s);
            c->cf.upload(cf.idx, s);
            c->ccf_host = cfdg::cell_cells_by_face(m);
            c->ccf_off.upload(c->ccf_host.off, s);
            c->ccf.upload(c->ccf_host.idx, s);
            std::vector<int> bfb(static_cast<std::size_t>(c->nbf));
            for (int i = 0; i < c->nbf; ++i) bfb[static_cast<std::size_t>(i)] = m.face_patch[static_cast<std::size_t>(c->ninner + i)];
            c->bface_bc.upload(bfb, s);
            c->sync();
        }
        if (num->gradient == CFDG_LEAST_SQUARES_NODE) {
            const cfdg::Csr st = cfdg::cell_cells_by_node(m);
            const std::vector<double> w = cfdg::lsq_coefficients(m, st);
            c->ccn_off.upload(st.off, s);
            c->ccn.upload(st.idx, s);
            c->lsq.upload(w, s);
            c->n_edges = static_cast<std::int64_t>(st.idx.size());
            c->sync();
        }
        {
            const cfdg::BsrPattern bp = cfdg::bsr_pattern(c->ccf_host, c->nc);
            c->h_row_ptr = bp.row_ptr;
            c->h_cols = bp.cols;
            c->h_diag = bp.diag;
            c->row_ptr.upload(bp.row_ptr, s);
            c->cols.upload(bp.cols, s);
            c->diag.upload(bp.diag, s);
            std::int32_t ncol = 0;
            const auto col = cfdg::greedy_colours(c->ccf_host, c->nc, ncol);
            const auto gf = cfdg::groups_from_labels(col, ncol, false);
            const auto gb = cfdg::groups_from_labels(col, ncol, true);
            c->h_gofs[0] = gf.off;
            c->h_gofs[1] = gb.off;
            c->grows[0].upload(gf.rows, s);
            c->grows[1].upload(gb.rows, s);
            c->sync();
        }
        const std::size_t nt = static_cast<std::size_t>(c->ntot), ncs = static_cast<std::size_t>(c->nc);
        c->U.alloc(5 * nt);
        c->q.alloc(5 * nt);
        c->grad.alloc(15 * nt);
        c->phi.alloc(5 * nt);
        c->res.alloc(5 * nt);
        c->lam.alloc(ncs);
        c->dt.alloc(ncs);
        c->alpha.alloc(ncs);
        c->U.zero(s);
        c->q.zero(s);
        c->grad.zero(s);
        c->phi.zero(s);
        c->res.zero(s);
        c->lam.zero(s);
        c->values.alloc(c->h_cols.size() * 25);
        c->values.zero(s);
        c->lu.alloc(ncs * 25);
        c->piv.alloc(ncs * 5);
        c->singular.alloc(1);
        for (DBuf<double>* v : {&c->b, &c->x, &c->r, &c->r0, &c->p, &c->v, &c->s, &c->t, &c->ph, &c->sh}) {
            v->alloc(5 * ncs);
            v->zero(s);
        }
        c->partial.alloc(static_cast<std::size_t>(cfdg::kReduceBlocks) * 5);
        c->dots.alloc(8);
        c->sync();
    });
    if (rc != CFDG_OK) {
        std::fprintf(stderr, "cfdg_create: %s\n", c->err.c_str());
        delete c;
        return rc;
    }
    *out = c;
    return CFDG_OK;
}

int cfdg_destroy(cfdg_ctx* ctx) {
    if (!ctx) return CFDG_ERR_ARG;
    delete ctx;
    return CFDG_OK;
}

const char* cfdg_last_error(const cfdg_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cfdg_sizes(const cfdg_ctx* ctx, int32_t* n_cells, int32_t* n_total, int64_t* nnzb) {
    if (!ctx) return CFDG_ERR_ARG;
    if (n_cells) *n_cells = ctx->nc;
    if (n_total) *n_total = ctx->ntot;
    if (nnzb) *nnzb = static_cast<int64_t>(ctx->h_cols.size());
    return CFDG_OK;
}

int cfdg_set_state(cfdg_ctx* c, const double* U, int64_t ld) {
    return guarded(c, [&] {
        if (!U || ld < c->nc) throw CfdgError(CFDG_ERR_ARG, "set_state: bad array");
        for (int v = 0; v < 5; ++v)
            hip_check(hipMemcpyAsync(c->U.get() + static_cast<std::size_t>(v) * static_cast<std::size_t>(c->ntot), U + v * ld,
                                     static_cast<std::size_t>(c->nc) * sizeof(double), hipMemcpyHostToDevice, c->stream),
                      "H2D");
        c->sync();
        c->have_residual = false;
        c->have_dt = false;
    });
}

int cfdg_get_state(cfdg_ctx* c, double* U, int64_t ld) {
    return guarded(c, [&] {
        if (!U || ld < c->nc) throw CfdgError(CFDG_ERR_ARG, "get_state: bad array");
        for (int v = 0; v < 5; ++v)
            hip_check(hipMemcpyAsync(U + v * ld, c->U.get() + static_cast<std::size_t>(v) * static_cast<std::size_t>(c->ntot),
                                     static_cast<std::size_t>(c->nc) * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                      "D2H");
        c->sync();
    });
}

int cfdg_init_uniform(cfdg_ctx* c, double rho, const double vel[3], double p) {
    return guarded(c, [&] {
        if (!vel) throw CfdgError(CFDG_ERR_ARG, "init_uniform: null velocity");
        // primitives_rhop_to_conserved (state_conversions.hpp:95-111), all n_total entries
        const double R = c->num.gas_constant, g = c->num.gamma;
        const double T = p / (rho * R);
        const double e = (R / (g - 1.0)) * T;
        const double ke = 0.5 * rho * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
        const double Uv[5] = {rho, rho * vel[0], rho * vel[1], rho * vel[2], rho * e + ke};
        std::vector<double> h(static_cast<std::size_t>(c->ntot) * 5);
        for (int v = 0; v < 5; ++v)
            std::fill_n(h.begin() + static_cast<std::ptrdiff_t>(v) * c->ntot, c->ntot, Uv[v]);
        hip_check(hipMemcpyAsync(c->U.get(), h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream), "H2D");
        c->sync();
        c->have_residual = false;
        c->have_dt = false;
    });
}

int cfdg_eval_residual(cfdg_ctx* c) {
    return guarded(c, [&] {
        eval_residual(*c);
        c->have_dt = false;
    });
}

int cfdg_get_primitives(cfdg_ctx* c, double* q, int64_t ld) {
    return guarded(c, [&] {
        if (!q || ld < c->ntot) throw CfdgError(CFDG_ERR_ARG, "get_primitives: bad array");
        for (int v = 0; v < 5; ++v)
            hip_check(hipMemcpyAsync(q + v * ld, c->q.get() + static_cast<std::size_t>(v) * static_cast<std::size_t>(c->ntot),
                                     static_cast<std::size_t>(c->ntot) * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                      "D2H");
        c->sync();
    });
}

int cfdg_get_gradients(cfdg_ctx* c, double* grad, int64_t ld, double* phi, int64_t ld_phi) {
    return guarded(c, [&] {
        if (grad) {
            if (ld < c->ntot) throw CfdgError(CFDG_ERR_ARG, "get_gradients: ld < n_total");
            for (int k = 0; k < 15; ++k)
                hip_check(hipMemcpyAsync(grad + k * ld, c->grad.get() + static_cast<std::size_t>(k) * static_cast<std::size_t>(c->ntot),
                                         static_cast<std::size_t>(c->ntot) * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                          "D2H");
        }
        if (phi) {
            if (ld_phi < c->nc) throw CfdgError(CFDG_ERR_ARG, "get_gradients: ld_phi < n_cells");
            for (int v = 0; v < 5; ++v)
                hip_check(hipMemcpyAsync(phi + v * ld_phi, c->phi.get() + static_cast<std::size_t>(v) * static_cast<std::size_t>(c->ntot),
                                         static_cast<std::size_t>(c->nc) * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                          "D2H");
        }
        c->sync();
    });
}

int cfdg_get_residual(cfdg_ctx* c, double* R, int64_t ld, double* lam) {
    return guarded(c, [&] {
        if (R) {
            if (ld < c->nc) throw CfdgError(CFDG_ERR_ARG, "get_residual: ld < n_cells");
            for (int v = 0; v < 5; ++v)
                hip_check(hipMemcpyAsync(R + v * ld, c->res.get() + static_cast<std::size_t>(v) * static_cast<std::size_t>(c->ntot),
                                         static_cast<std::size_t>(c->nc) * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                          "D2H");
        }
        if (lam)
            hip_check(hipMemcpyAsync(lam, c->lam.get(), static_cast<std::size_t>(c->nc) * sizeof(double), hipMemcpyDeviceToHost,
                                     c->stream),
                      "D2H");
        c->sync();
    });
}

int cfdg_residual_norms(cfdg_ctx* c, double l2[5]) {
    return guarded(c, [&] {
        if (!l2) throw CfdgError(CFDG_ERR_ARG, "null output");
        residual_norms(*c, l2);
    });
}

int cfdg_compute_dt(cfdg_ctx* c) { return guarded(c, [&] { compute_dt(*c); }); }

int cfdg_get_dt(cfdg_ctx* c, double* dt, double* alpha) {
    return guarded(c, [&] {
        if (dt)
            hip_check(hipMemcpyAsync(dt, c->dt.get(), static_cast<std::size_t>(c->nc) * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                      "D2H");
        if (alpha)
            hip_check(hipMemcpyAsync(alpha, c->alpha.get(), static_cast<std::size_t>(c->nc) * sizeof(double), hipMemcpyDeviceToHost,
                                     c->stream),
                      "D2H");
        c->sync();
    });
}

int cfdg_assemble_jacobian(cfdg_ctx* c) { return guarded(c, [&] { assemble_jacobian(*c); }); }

int cfdg_get_bsr_pattern(cfdg_ctx* c, int32_t* row_ptr, int32_t* cols, int32_t* diag_idx) {
    return guarded(c, [&] {
        if (row_ptr) std::copy(c->h_row_ptr.begin(), c->h_row_ptr.end(), row_ptr);
        if (cols) std::copy(c->h_cols.begin(), c->h_cols.end(), cols);
        if (diag_idx) std::copy(c->h_diag.begin(), c->h_diag.end(), diag_idx);
    });
}

int cfdg_get_bsr_values(cfdg_ctx* c, double* values) {
    return guarded(c, [&] {
        if (!values) throw CfdgError(CFDG_ERR_ARG, "null output");
        hip_check(hipMemcpyAsync(values, c->values.get(), c->values.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                  "D2H");
        c->sync();
    });
}

int cfdg_set_bsr_values(cfdg_ctx* c, const double* values) {
    return guarded(c, [&] {
        if (!values) throw CfdgError(CFDG_ERR_ARG, "null input");
        hip_check(hipMemcpyAsync(c->values.get(), values, c->values.size() * sizeof(double), hipMemcpyHostToDevice, c->stream),
                  "H2D");
        c->sync();
    });
}

int cfdg_linear_solve(cfdg_ctx* c, const cfdg_linparams* lp, const double* b, const double* x0, double* x,
                      cfdg_iter_result* res) {
    return guarded(c, [&] {
        if (!lp) throw CfdgError(CFDG_ERR_ARG, "null linear parameters");
        linear_solve(*c, *lp, b, x0, x, res);
    });
}

int cfdg_sgs_apply(cfdg_ctx* c, int32_t sgs_order, const double* r, double* z) {
    return guarded(c, [&] {
        if (!r || !z) throw CfdgError(CFDG_ERR_ARG, "null vector");
        if (sgs_order != CFDG_SGS_MULTICOLOUR && sgs_order != CFDG_SGS_SEQUENTIAL)
            throw CfdgError(CFDG_ERR_CONFIG, "unknown SGS order");
        const std::size_t bytes = 5ULL * static_cast<std::size_t>(c->nc) * sizeof(double);
        Precond P;
        P.kind = CFDG_PRECOND_SGS;
        P.order = sgs_order;
        precond_setup(*c, P);
        hip_check(hipMemcpyAsync(c->s.get(), r, bytes, hipMemcpyHostToDevice, c->stream), "H2D");
        precond_apply(*c, P, c->s.get(), c->sh.get());
        hip_check(hipMemcpyAsync(z, c->sh.get(), bytes, hipMemcpyDeviceToHost, c->stream), "D2H");
        c->check_launch();
        c->sync();
    });
}

int cfdg_spmv(cfdg_ctx* c, const double* x, double* y) {
    return guarded(c, [&] {
        if (!x || !y) throw CfdgError(CFDG_ERR_ARG, "null vector");
        const std::size_t bytes = 5ULL * static_cast<std::size_t>(c->nc) * sizeof(double);
        hip_check(hipMemcpyAsync(c->s.get(), x, bytes, hipMemcpyHostToDevice, c->stream), "H2D");
        cfdg::launch_spmv(c->bargs(), c->values.get(), c->s.get(), c->t.get(), c->stream);
        hip_check(hipMemcpyAsync(y, c->t.get(), bytes, hipMemcpyDeviceToHost, c->stream), "D2H");
        c->check_launch();
        c->sync();
    });
}

int cfdg_implicit_step(cfdg_ctx* c, const cfdg_linparams* lp, cfdg_step_result* out) {
    return guarded(c, [&] {
        if (!lp) throw CfdgError(CFDG_ERR_ARG, "null linear parameters");
        eval_residual(*c);
        compute_dt(*c);
        assemble_jacobian(*c);
        cfdg_iter_result lr{};
        linear_solve(*c, *lp, nullptr, nullptr, nullptr, &lr);
        double l2[5];
        residual_norms(*c, l2);
        cfdg::launch_state_update(c->nc, c->ntot, c->x.get(), c->U.get(), c->stream);
        c->check_launch();
        c->sync();
        c->have_residual = false;
        c->have_dt = false;
        if (out) {
            for (int v = 0; v < 5; ++v) out->l2[v] = l2[v];
            out->linear = lr;
        }
    });
}

int cfdg_stage_times(cfdg_ctx* c, double ms[6]) {
    return guarded(c, [&] {
        if (!ms) throw CfdgError(CFDG_ERR_ARG, "null output");
        for (int i = 0; i < 6; ++i) ms[i] = c->stage_ms[i];
    });
}

int cfdg_time_flux_kernel(cfdg_ctx* c, int32_t reps, double* ms_per_launch) {
    return guarded(c, [&] {
        if (reps < 1 || !ms_per_launch) throw CfdgError(CFDG_ERR_ARG, "reps >= 1 and an output are required");
        if (!c->have_residual) throw CfdgError(CFDG_ERR_STATE, "time_flux_kernel needs an evaluated residual");
        const auto m = c->margs();
        const auto f = c->fargs();
        hip_check(hipEventRecord(c->ev[6], c->stream), "event");
        for (int i = 0; i < reps; ++i) cfdg::launch_flux(m, f, c->nargs, c->stream);
        hip_check(hipEventRecord(c->ev[7], c->stream), "event");
        c->check_launch();
        c->sync();
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, c->ev[6], c->ev[7]), "elapsed");
        *ms_per_launch = static_cast<double>(ms) / reps;
    });
}

}  // extern "C"
