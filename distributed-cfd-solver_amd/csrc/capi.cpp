// C ABI of the solver context (include/cfdg.h): set-up, device memory, and
// the orchestration of the residual, Jacobian and BiCGSTAB stages on one HIP
// stream.  Every entry point catches exceptions and returns a status code.
//
// Internally cells live in the device (tile) order of setup.hpp; every array
// crossing the ABI is in mesh order and is permuted here, on the host.
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

#include "capi_internal.hpp"
#include "kernels.hpp"
#include "kernels_linalg.hpp"
#include "mesh.hpp"
#include "setup.hpp"

namespace {

using cfdg::CfdgError;
using i32 = std::int32_t;
inline std::size_t z(long long v) { return static_cast<std::size_t>(v); }

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

// Resolved boundary parameters (bc_manager.hpp:235-352 builders, resolve_inflow :37-111).
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

// CSR list of mesh-numbered cells re-expressed for device rows: row r holds
// the list of mesh cell new2old[r], entries mapped through map (or kept when
// map is null), order unchanged.
cfdg::Csr to_device_rows(const cfdg::Csr& in, const std::vector<i32>& new2old, const std::vector<i32>* map) {
    cfdg::Csr out;
    const std::size_t n = new2old.size();
    out.off.assign(n + 1, 0);
    for (std::size_t r = 0; r < n; ++r) {
        const i32 c = new2old[r];
        out.off[r + 1] = out.off[r] + (in.off[z(c) + 1] - in.off[z(c)]);
    }
    out.idx.resize(z(out.off.back()));
    for (std::size_t r = 0; r < n; ++r) {
        const i32 c = new2old[r];
        i32 w = out.off[r];
        for (i32 e = in.off[z(c)]; e < in.off[z(c) + 1]; ++e) out.idx[z(w++)] = map ? (*map)[z(in.idx[z(e)])] : in.idx[z(e)];
    }
    return out;
}

}  // namespace

struct cfdg_ctx {
    std::string err;
    hipStream_t stream = nullptr;
    int nc = 0, nf = 0, ninner = 0, nbf = 0, ntot = 0, npatch = 0, ntiles = 0;
    cfdg_numerics num{};
    cfdg::NumArgs nargs{};
    cfdg::BcTable bcs{};
    std::vector<i32> new2old, old2new;

    // mesh (device)
    DBuf<double> frec;
    DBuf<int> cf_off, cf, cf_slot, ccf_off, ccf, ccn_off, ccn, bface_bc, tile_off, tile_face_off, tile_faces;
    DBuf<int> tile_halo_off, tile_halo;
    DBuf<std::uint16_t> ccn_slot, ccf_slot;
    // fields (device order)
    DBuf<double> U, crec, res, lam, dt, alpha;
    bool have_residual = false;
    bool have_dt = false;
    // BSR: mesh-order pattern for the ABI, device-order pattern for the kernels
    std::vector<i32> h_row_ptr, h_cols, h_diag;
    std::vector<long long> blk_dev;  // mesh block k -> device block
    DBuf<int> row_ptr, cols, diag;
    DBuf<double> values, lu;
    DBuf<std::uint8_t> piv;
    DBuf<int> singular;
    // SGS row groups (device rows): 0 multicolour fwd, 1 multicolour bwd, 2 sequential fwd, 3 sequential bwd
    std::vector<i32> h_gofs[4];
    DBuf<int> grows[4];
    cfdg::Csr ccf_mesh;
    bool have_seq_groups = false;
    // Krylov work vectors (AoS, 5 nc, device order)
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
        a.ntiles = ntiles;
        a.frec = frec.get();
        a.cf_off = cf_off.get();
        a.cf = cf.get();
        a.cf_slot = cf_slot.get();
        a.ccf_off = ccf_off.get();
        a.ccf = ccf.get();
        a.ccn_off = ccn_off.get();
        a.ccn = ccn.get();
        a.bface_bc = bface_bc.get();
        a.tile_off = tile_off.get();
        a.tile_face_off = tile_face_off.get();
        a.tile_faces = tile_faces.get();
        a.tile_halo_off = tile_halo_off.get();
        a.tile_halo = tile_halo.get();
        a.ccn_slot = ccn_slot.get();
        a.ccf_slot = ccf_slot.get();
        return a;
    }
    cfdg::FieldArgs fargs() const {
        cfdg::FieldArgs f{};
        f.U = U.get();
        f.crec = crec.get();
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
    template <typename T>
    void d2h(T* dst, const T* src, std::size_t n) {
        if (n) hip_check(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, stream), "D2H");
    }
    template <typename T>
    void h2d(T* dst, const T* src, std::size_t n) {
        if (n) hip_check(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, stream), "H2D");
    }
    // AoS block vector (5 per cell): mesh order <-> device order
    std::vector<double> vec_to_device(const double* x_mesh) const {
        std::vector<double> d(z(nc) * 5);
        for (int r = 0; r < nc; ++r)
            for (int k = 0; k < 5; ++k) d[z(r) * 5 + z(k)] = x_mesh[z(new2old[z(r)]) * 5 + z(k)];
        return d;
    }
    void vec_to_mesh(const std::vector<double>& d, double* x_mesh) const {
        for (int r = 0; r < nc; ++r)
            for (int k = 0; k < 5; ++k) x_mesh[z(new2old[z(r)]) * 5 + z(k)] = d[z(r) * 5 + z(k)];
    }
    std::vector<double> download_records() {
        std::vector<double> h(z(ntot) * cfdg::kRec);
        d2h(h.data(), crec.get(), h.size());
        sync();
        return h;
    }

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

void upload_groups(cfdg_ctx& c, int slot, const std::vector<i32>& label_mesh, i32 n_labels, bool reverse) {
    const cfdg::RowGroups g = cfdg::groups_from_labels(label_mesh, n_labels, reverse);
    std::vector<i32> rows(g.rows.size());
    for (std::size_t i = 0; i < rows.size(); ++i) rows[i] = c.old2new[z(g.rows[i])];
    c.h_gofs[slot] = g.off;
    c.grows[slot].upload(rows, c.stream);
}

void build_seq_groups(cfdg_ctx& c) {
    if (c.have_seq_groups) return;
    i32 nfw = 0, nbw = 0;
    const auto lf = cfdg::forward_levels(c.ccf_mesh, c.nc, nfw);
    const auto lb = cfdg::backward_levels(c.ccf_mesh, c.nc, nbw);
    upload_groups(c, 2, lf, nfw, false);
    upload_groups(c, 3, lb, nbw, false);
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
    for (int k = 0; k < 4; ++k) {
        hip_check(hipEventElapsedTime(&ms, c.ev[k], c.ev[k + 1]), "elapsed");
        c.stage_ms[k] = ms;
    }
    c.have_residual = true;
}

void compute_dt(cfdg_ctx& c) {
    if (!c.have_residual) throw CfdgError(CFDG_ERR_STATE, "compute_dt before eval_residual");
    cfdg::launch_dt(c.margs(), c.fargs(), c.lam.get(), c.num.cfl, c.dt.get(), c.alpha.get(), c.stream);
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
    c.d2h(c.h_dots, c.dots.get(), static_cast<std::size_t>(n));
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
        for (std::size_t k = 0; k + 1 < off.size(); ++k)
            cfdg::launch_sgs_group(c.grows[gi].get() + off[k], off[k + 1] - off[k], b, c.values.get(), c.lu.get(),
                                   c.piv.get(), r, z, c.stream);
    }
}

void precond_setup(cfdg_ctx& c, const Precond& P) {
    if (P.kind != CFDG_PRECOND_SGS) return;
    if (P.order == CFDG_SGS_SEQUENTIAL) build_seq_groups(c);
    c.singular.zero(c.stream);
    cfdg::launch_lu_factor(c.nc, c.diag.get(), c.values.get(), c.lu.get(), c.piv.get(), c.singular.get(), c.stream);
    c.check_launch();
    int sing = 0;
    c.d2h(&sing, c.singular.get(), 1);
    c.sync();
    if (sing != 0)
        throw CfdgError(CFDG_ERR_SINGULAR,
                        "SGS: singular diagonal block at row " + std::to_string(c.new2old[z(sing - 1)]));
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
    const std::size_t n5 = 5ULL * static_cast<std::size_t>(c.nc);
    if (hb) {
        const auto d = c.vec_to_device(hb);
        c.h2d(c.b.get(), d.data(), n5);
        c.sync();
    } else {
        if (!c.have_residual) throw CfdgError(CFDG_ERR_STATE, "implicit right-hand side needs a residual");
        cfdg::launch_rhs(c.nc, c.nc, c.res.get(), c.b.get(), c.stream);
    }
    if (hx0) {
        const auto d = c.vec_to_device(hx0);
        c.h2d(c.x.get(), d.data(), n5);
        c.sync();
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
    c.sync();
    if (hx) {
        std::vector<double> d(n5);
        c.d2h(d.data(), c.x.get(), n5);
        c.sync();
        c.vec_to_mesh(d, hx);
    }
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, c.ev[5], c.ev[6]), "elapsed");
    c.stage_ms[5] = ms;
    if (out) *out = r;
}

void residual_norms(cfdg_ctx& c, double l2[5]) {
    if (!c.have_residual) throw CfdgError(CFDG_ERR_STATE, "residual norms before eval_residual");
    cfdg::launch_res_sq(c.nc, c.nc, c.res.get(), c.partial.get(), c.dots.get(), c.stream);
    fetch_dots(c, 5);
    const double scale = 1.0 / static_cast<double>(std::max(c.nc, 1));
    for (int v = 0; v < 5; ++v) l2[v] = std::sqrt(c.h_dots[v] * scale);
}

void check_numerics(const cfdg_numerics* num) {
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
}

void setup_context(cfdg_ctx& c, const cfdg::Mesh& m, const cfdg_numerics* num, const cfdg_bc* bcs, int32_t n_bcs) {
    if (n_bcs != static_cast<int32_t>(m.patch_names.size()))
        throw CfdgError(CFDG_ERR_ARG, "one boundary condition per mesh patch is required");
    if (n_bcs > cfdg::kMaxPatches) throw CfdgError(CFDG_ERR_ARG, "too many patches");
    check_numerics(num);

    int ndev = 0;
    hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if (ndev < 1) throw CfdgError(CFDG_ERR_HIP, "no HIP device");
    hip_check(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& e : c.ev) hip_check(hipEventCreate(&e), "hipEventCreate");
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&c.h_dots), 8 * sizeof(double)), "hipHostMalloc");

    c.num = *num;
    c.nc = m.n_cells;
    c.nf = m.n_faces;
    c.ninner = m.n_inner_faces;
    c.nbf = m.n_faces - m.n_inner_faces;
    c.ntot = c.nc + c.nbf;
    c.npatch = n_bcs;
    c.nargs.gas = cfdg::Gas{num->gamma, num->gas_constant, num->prandtl, num->prandtl_turb};
    c.nargs.viscous = num->flow_model == CFDG_VISCOUS_FLOW;
    c.nargs.muscl = num->reconstruction == CFDG_MUSCL;
    c.nargs.limiter = num->limiter;
    c.nargs.gradient = num->gradient;
    c.nargs.k3 = num->venkat_k * num->venkat_k * num->venkat_k;
    for (int32_t p = 0; p < n_bcs; ++p) c.bcs.bc[p] = resolve_bc(bcs[p], num->gamma, num->gas_constant);
    hipStream_t s = c.stream;

    // mesh-numbered connectivity, device tiling and permutation
    const cfdg::Csr cf_mesh = cfdg::cell_faces(m);
    c.ccf_mesh = cfdg::cell_cells_by_face(m);
    const cfdg::Tiling tl = cfdg::tile_cells(m, c.ccf_mesh, cf_mesh, cfdg::kTileCells, cfdg::kTileFaces);
    c.new2old = tl.new2old;
    c.old2new = tl.old2new;
    c.ntiles = static_cast<int>(tl.tile_off.size()) - 1;

    // face records in device numbering
    std::vector<i32> owner_dev(z(c.nf)), neigh_dev(z(c.nf));
    {
        std::vector<double> dx, dy, dz, di;
        cfdg::face_distances(m, dx, dy, dz, di);
        std::vector<double> fr(z(c.nf) * cfdg::kFRec, 0.0);
        for (int f = 0; f < c.nf; ++f) {
            double* r = fr.data() + z(f) * cfdg::kFRec;
            r[cfdg::kFN] = m.face_nx[z(f)];
            r[cfdg::kFN + 1] = m.face_ny[z(f)];
            r[cfdg::kFN + 2] = m.face_nz[z(f)];
            r[cfdg::kFA] = m.face_area[z(f)];
            r[cfdg::kFC] = m.face_cx[z(f)];
            r[cfdg::kFC + 1] = m.face_cy[z(f)];
            r[cfdg::kFC + 2] = m.face_cz[z(f)];
            r[cfdg::kFD] = dx[z(f)];
            r[cfdg::kFD + 1] = dy[z(f)];
            r[cfdg::kFD + 2] = dz[z(f)];
            r[cfdg::kFDinv] = di[z(f)];
            owner_dev[z(f)] = c.old2new[z(m.face_owner[z(f)])];
            neigh_dev[z(f)] = (m.face_neigh[z(f)] >= 0) ? c.old2new[z(m.face_neigh[z(f)])] : -1;
            i32 cells[2] = {owner_dev[z(f)], neigh_dev[z(f)]};
            std::memcpy(r + cfdg::kFCells, cells, sizeof(cells));
        }
        c.frec.upload(fr, s);
        c.sync();
    }
    {
        const cfdg::Csr cf_dev = to_device_rows(cf_mesh, c.new2old, nullptr);
        const cfdg::TileFaces tf = cfdg::tile_faces(tl, cf_dev, owner_dev);
        for (int k = 0; k < c.ntiles; ++k) {
            if (tf.tile_face_off[z(k) + 1] - tf.tile_face_off[z(k)] > cfdg::kTileFaces ||
                tl.tile_off[z(k) + 1] - tl.tile_off[z(k)] > cfdg::kTileCells)
                throw CfdgError(CFDG_ERR_MESH, "tile exceeds the flux kernel's LDS capacity");
        }
        c.cf_off.upload(cf_dev.off, s);
        c.cf.upload(cf_dev.idx, s);
        c.cf_slot.upload(tf.cell_slot, s);
        c.tile_off.upload(tl.tile_off, s);
        c.tile_face_off.upload(tf.tile_face_off, s);
        c.tile_faces.upload(tf.faces, s);
        const cfdg::Csr ccf_dev = to_device_rows(c.ccf_mesh, c.new2old, &c.old2new);
        c.ccf_off.upload(ccf_dev.off, s);
        c.ccf.upload(ccf_dev.idx, s);
        std::vector<int> bfb(z(c.nbf));
        for (int i = 0; i < c.nbf; ++i) bfb[z(i)] = m.face_patch[z(c.ninner + i)];
        c.bface_bc.upload(bfb, s);
        c.sync();
    }
    if (num->gradient == CFDG_LEAST_SQUARES_NODE) {
        const cfdg::Csr st = to_device_rows(cfdg::cell_cells_by_node(m), c.new2old, &c.old2new);
        c.ccn_off.upload(st.off, s);
        c.ccn.upload(st.idx, s);
        // LDS-staged stencils when every tile's cells + halo fit the kernel's slots
        const cfdg::Csr ccf_dev = to_device_rows(c.ccf_mesh, c.new2old, &c.old2new);
        const cfdg::TileStencil ts = cfdg::tile_stencil(tl, st, ccf_dev);
        if (ts.max_slots <= cfdg::kGradSlots) {
            c.tile_halo_off.upload(ts.halo_off, s);
            c.tile_halo.upload(ts.halo, s);
            c.ccn_slot.upload(ts.ccn_slot, s);
            c.ccf_slot.upload(ts.ccf_slot, s);
        }
        c.sync();
    }
    // cell records: static centroid and volume; everything else zero
    {
        std::vector<double> cr(z(c.ntot) * cfdg::kRec, 0.0);
        for (int r = 0; r < c.nc; ++r) {
            const std::size_t o = z(c.new2old[z(r)]);
            double* p = cr.data() + z(r) * cfdg::kRec;
            p[cfdg::kRC] = m.cell_cx[o];
            p[cfdg::kRC + 1] = m.cell_cy[o];
            p[cfdg::kRC + 2] = m.cell_cz[o];
            p[cfdg::kRVol] = m.cell_vol[o];
        }
        c.crec.upload(cr, s);
        c.sync();
    }
    // BSR: the mesh-order pattern (bsr_helpers.hpp:16-59) and its device rows
    {
        const cfdg::BsrPattern bp = cfdg::bsr_pattern(c.ccf_mesh, c.nc);
        c.h_row_ptr = bp.row_ptr;
        c.h_cols = bp.cols;
        c.h_diag = bp.diag;
        std::vector<i32> rp(z(c.nc) + 1, 0), cols(bp.cols.size()), dg(z(c.nc));
        for (int r = 0; r < c.nc; ++r) {
            const i32 o = c.new2old[z(r)];
            rp[z(r) + 1] = rp[z(r)] + (bp.row_ptr[z(o) + 1] - bp.row_ptr[z(o)]);
        }
        c.blk_dev.assign(bp.cols.size(), 0);
        for (int r = 0; r < c.nc; ++r) {
            const i32 o = c.new2old[z(r)];
            for (i32 k = bp.row_ptr[z(o)]; k < bp.row_ptr[z(o) + 1]; ++k) {
                const i32 kd = rp[z(r)] + (k - bp.row_ptr[z(o)]);
                cols[z(kd)] = c.old2new[z(bp.cols[z(k)])];
                c.blk_dev[z(k)] = kd;
            }
            dg[z(r)] = rp[z(r)] + (bp.diag[z(o)] - bp.row_ptr[z(o)]);
        }
        c.row_ptr.upload(rp, s);
        c.cols.upload(cols, s);
        c.diag.upload(dg, s);
        i32 ncol = 0;
        const auto col = cfdg::greedy_colours(c.ccf_mesh, c.nc, ncol);
        upload_groups(c, 0, col, ncol, false);
        upload_groups(c, 1, col, ncol, true);
        c.sync();
    }
    const std::size_t ncs = z(c.nc);
    c.U.alloc(5 * ncs);
    c.res.alloc(5 * ncs);
    c.lam.alloc(ncs);
    c.dt.alloc(ncs);
    c.alpha.alloc(ncs);
    c.U.zero(s);
    c.res.zero(s);
    c.lam.zero(s);
    c.values.alloc(c.h_cols.size() * 25);
    c.values.zero(s);
    c.lu.alloc(ncs * 25);
    c.piv.alloc(ncs * 5);
    c.singular.alloc(1);
    for (DBuf<double>* v : {&c.b, &c.x, &c.r, &c.r0, &c.p, &c.v, &c.s, &c.t, &c.ph, &c.sh}) {
        v->alloc(5 * ncs);
        v->zero(s);
    }
    c.partial.alloc(static_cast<std::size_t>(cfdg::kReduceBlocks) * 5);
    c.dots.alloc(8);
    c.sync();
}

}  // namespace

extern "C" {

int cfdg_create(const cfdg_mesh* mesh, const cfdg_numerics* num, const cfdg_bc* bcs, int32_t n_bcs, cfdg_ctx** out) {
    if (!mesh || !num || !out || (n_bcs > 0 && !bcs)) return CFDG_ERR_ARG;
    *out = nullptr;
    auto* c = new cfdg_ctx();
    const int rc = guarded(c, [&] { setup_context(*c, mesh->m, num, bcs, n_bcs); });
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
        std::vector<double> h(5 * z(c->nc));
        for (int v = 0; v < 5; ++v)
            for (int r = 0; r < c->nc; ++r) h[z(v) * z(c->nc) + z(r)] = U[v * ld + c->new2old[z(r)]];
        c->h2d(c->U.get(), h.data(), h.size());
        c->sync();
        c->have_residual = false;
        c->have_dt = false;
    });
}

int cfdg_get_state(cfdg_ctx* c, double* U, int64_t ld) {
    return guarded(c, [&] {
        if (!U || ld < c->nc) throw CfdgError(CFDG_ERR_ARG, "get_state: bad array");
        std::vector<double> h(5 * z(c->nc));
        c->d2h(h.data(), c->U.get(), h.size());
        c->sync();
        for (int v = 0; v < 5; ++v)
            for (int r = 0; r < c->nc; ++r) U[v * ld + c->new2old[z(r)]] = h[z(v) * z(c->nc) + z(r)];
    });
}

int cfdg_init_uniform(cfdg_ctx* c, double rho, const double vel[3], double p) {
    return guarded(c, [&] {
        if (!vel) throw CfdgError(CFDG_ERR_ARG, "init_uniform: null velocity");
        // primitives_rhop_to_conserved (state_conversions.hpp:95-111)
        const double R = c->num.gas_constant, g = c->num.gamma;
        const double T = p / (rho * R);
        const double e = (R / (g - 1.0)) * T;
        const double ke = 0.5 * rho * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
        const double Uv[5] = {rho, rho * vel[0], rho * vel[1], rho * vel[2], rho * e + ke};
        std::vector<double> h(z(c->nc) * 5);
        for (int v = 0; v < 5; ++v) std::fill_n(h.begin() + static_cast<std::ptrdiff_t>(v) * c->nc, c->nc, Uv[v]);
        c->h2d(c->U.get(), h.data(), h.size());
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
        const auto h = c->download_records();
        for (int r = 0; r < c->ntot; ++r) {
            const long long o = (r < c->nc) ? c->new2old[z(r)] : r;
            for (int v = 0; v < 5; ++v) q[v * ld + o] = h[z(r) * cfdg::kRec + z(cfdg::kRQ + v)];
        }
    });
}

int cfdg_get_gradients(cfdg_ctx* c, double* grad, int64_t ld, double* phi, int64_t ld_phi) {
    return guarded(c, [&] {
        if (grad && ld < c->ntot) throw CfdgError(CFDG_ERR_ARG, "get_gradients: ld < n_total");
        if (phi && ld_phi < c->nc) throw CfdgError(CFDG_ERR_ARG, "get_gradients: ld_phi < n_cells");
        const auto h = c->download_records();
        for (int r = 0; r < c->ntot; ++r) {
            const long long o = (r < c->nc) ? c->new2old[z(r)] : r;
            const double* p = h.data() + z(r) * cfdg::kRec;
            if (grad)
                for (int k = 0; k < 15; ++k) grad[k * ld + o] = p[cfdg::kRG + k];
            if (phi && r < c->nc)
                for (int v = 0; v < 5; ++v) phi[v * ld_phi + o] = p[cfdg::kRPhi + v];
        }
    });
}

int cfdg_get_residual(cfdg_ctx* c, double* R, int64_t ld, double* lam) {
    return guarded(c, [&] {
        if (R && ld < c->nc) throw CfdgError(CFDG_ERR_ARG, "get_residual: ld < n_cells");
        std::vector<double> h(5 * z(c->nc)), l(z(c->nc));
        c->d2h(h.data(), c->res.get(), h.size());
        c->d2h(l.data(), c->lam.get(), l.size());
        c->sync();
        for (int r = 0; r < c->nc; ++r) {
            const long long o = c->new2old[z(r)];
            if (R)
                for (int v = 0; v < 5; ++v) R[v * ld + o] = h[z(v) * z(c->nc) + z(r)];
            if (lam) lam[o] = l[z(r)];
        }
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
        std::vector<double> a(z(c->nc)), b(z(c->nc));
        c->d2h(a.data(), c->dt.get(), a.size());
        c->d2h(b.data(), c->alpha.get(), b.size());
        c->sync();
        for (int r = 0; r < c->nc; ++r) {
            if (dt) dt[c->new2old[z(r)]] = a[z(r)];
            if (alpha) alpha[c->new2old[z(r)]] = b[z(r)];
        }
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
        std::vector<double> h(c->values.size());
        c->d2h(h.data(), c->values.get(), h.size());
        c->sync();
        for (std::size_t k = 0; k < c->blk_dev.size(); ++k)
            std::memcpy(values + k * 25, h.data() + z(c->blk_dev[k]) * 25, 25 * sizeof(double));
    });
}

int cfdg_set_bsr_values(cfdg_ctx* c, const double* values) {
    return guarded(c, [&] {
        if (!values) throw CfdgError(CFDG_ERR_ARG, "null input");
        std::vector<double> h(c->values.size());
        for (std::size_t k = 0; k < c->blk_dev.size(); ++k)
            std::memcpy(h.data() + z(c->blk_dev[k]) * 25, values + k * 25, 25 * sizeof(double));
        c->h2d(c->values.get(), h.data(), h.size());
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

int cfdg_sgs_apply(cfdg_ctx* c, int32_t sgs_order, const double* r, double* zout) {
    return guarded(c, [&] {
        if (!r || !zout) throw CfdgError(CFDG_ERR_ARG, "null vector");
        if (sgs_order != CFDG_SGS_MULTICOLOUR && sgs_order != CFDG_SGS_SEQUENTIAL)
            throw CfdgError(CFDG_ERR_CONFIG, "unknown SGS order");
        const std::size_t n5 = 5ULL * z(c->nc);
        Precond P;
        P.kind = CFDG_PRECOND_SGS;
        P.order = sgs_order;
        precond_setup(*c, P);
        const auto d = c->vec_to_device(r);
        c->h2d(c->s.get(), d.data(), n5);
        precond_apply(*c, P, c->s.get(), c->sh.get());
        std::vector<double> h(n5);
        c->d2h(h.data(), c->sh.get(), n5);
        c->check_launch();
        c->sync();
        c->vec_to_mesh(h, zout);
    });
}

int cfdg_spmv(cfdg_ctx* c, const double* x, double* y) {
    return guarded(c, [&] {
        if (!x || !y) throw CfdgError(CFDG_ERR_ARG, "null vector");
        const std::size_t n5 = 5ULL * z(c->nc);
        const auto d = c->vec_to_device(x);
        c->h2d(c->s.get(), d.data(), n5);
        cfdg::launch_spmv(c->bargs(), c->values.get(), c->s.get(), c->t.get(), c->stream);
        std::vector<double> h(n5);
        c->d2h(h.data(), c->t.get(), n5);
        c->check_launch();
        c->sync();
        c->vec_to_mesh(h, y);
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
        cfdg::launch_state_update(c->nc, c->nc, c->x.get(), c->U.get(), c->stream);
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
