// C ABI of the solver context (include/cfdg.h): set-up, device memory, and
// the orchestration of the residual, Jacobian and BiCGSTAB stages on one HIP
// stream.  Every entry point catches exceptions and returns a status code.
//
// Internally cells live in the device (tile) order of setup.hpp; every array
// crossing the ABI is in mesh order and is permuted here, on the host.
#include "cfdg.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "ctx.hpp"
#include "kernels.hpp"
#include "kernels_linalg.hpp"
#include "mesh.hpp"
#include "setup.hpp"

namespace {

using cfdg::CfdgError;
using cfdg::DBuf;
using cfdg::hip_check;
using cfdg::z;
using i32 = std::int32_t;

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

// Forward-sweep flags of every block (device order) from the rows' group
// labels (mesh order): 1 when the column's group precedes the row's; and of
// every device row: 1 when it lies in the first group (label 0).
void upload_lower_flags(cfdg_ctx& c, int slot, const std::vector<i32>& label_mesh) {
    std::vector<std::uint8_t> f(c.h_cols.size(), 0), first(z(c.nc), 0);
    for (i32 o = 0; o < c.nc; ++o) {
        first[z(c.old2new[z(o)])] = label_mesh[z(o)] == 0 ? 1 : 0;
        for (i32 k = c.h_row_ptr[z(o)]; k < c.h_row_ptr[z(o) + 1]; ++k)
            f[z(c.blk_dev[z(k)])] = label_mesh[z(c.h_cols[z(k)])] < label_mesh[z(o)] ? 1 : 0;
    }
    c.blk_lower[slot].upload(f, c.stream);
    c.row_first[slot].upload(first, c.stream);
    c.sync();  // f and first are released on return
}

void build_seq_groups(cfdg_ctx& c) {
    if (c.have_seq_groups) return;
    i32 nfw = 0, nbw = 0;
    const auto lf = cfdg::forward_levels(c.ccf_mesh, c.nc, nfw);
    const auto lb = cfdg::backward_levels(c.ccf_mesh, c.nc, nbw);
    upload_groups(c, 2, lf, nfw, false);
    upload_groups(c, 3, lb, nbw, false);
    upload_lower_flags(c, 1, lf);
    c.sync();
    c.have_seq_groups = true;
}

void eval_residual(cfdg_ctx& c) {
    c.have_features = false;
    const auto m = c.margs();
    const auto f = c.fargs();
    hip_check(hipEventRecord(c.ev[0], c.stream), "event");
    cfdg::launch_primitives(m, f, c.nargs, c.stream);
    cfdg::launch_ghosts(m, f, c.nargs, c.bcs, c.stream);
    // module ghosts, then density / nu_tilde / eddy viscosity (solver.hpp:195-204)
    if (c.sa_on()) {
        cfdg::launch_sa_prepare(m, f, c.nargs, c.sargs(), c.stream);
        c.sa_evaluated = true;
    }
    hip_check(hipEventRecord(c.ev[1], c.stream), "event");
    const bool need_grad = c.nargs.muscl || c.nargs.viscous;
    if (need_grad) {
        const cfdg::SaArgs sa = c.sargs();
        const bool sa_done = cfdg::launch_gradient_limiter(m, f, c.nargs, c.stream, c.sa_on() ? &sa : nullptr);
        hip_check(hipEventRecord(c.ev[2], c.stream), "event");
        cfdg::launch_ghost_gradients(m, f, c.nargs, c.bcs, c.stream);
        if (c.sa_on() && !sa_done) cfdg::launch_sa_gradient(m, f, c.nargs, sa, c.stream);
    } else {
        hip_check(hipEventRecord(c.ev[2], c.stream), "event");
    }
    hip_check(hipEventRecord(c.ev[3], c.stream), "event");
    cfdg::launch_flux(m, f, c.nargs, c.stream);
    // module face sweep on the recorded mass flux, then its sources (solver.hpp:233-241)
    if (c.sa_on()) cfdg::launch_sa_residual(m, f, c.nargs, c.sargs(), c.stream);
    // BDF pseudo-source of dual time stepping (solver.hpp:243-246)
    if (c.bdf_order > 0)
        cfdg::launch_unsteady_source(m, f, c.bdf_cur, 1.0 / c.phys_dt, c.Un.get(), c.Unm1.get(), c.stream);
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
    const double unsteady = (c.bdf_order > 0) ? ((c.bdf_cur == 2) ? 1.5 : 1.0) / c.phys_dt : 0.0;
    cfdg::launch_dt(c.margs(), c.fargs(), c.lam.get(), c.num.cfl, unsteady, c.dt.get(), c.alpha.get(), c.stream);
    c.check_launch();
    c.sync();
    c.have_dt = true;
}

void assemble_jacobian(cfdg_ctx& c) {
    if (!c.have_dt) throw CfdgError(CFDG_ERR_STATE, "assemble_jacobian before compute_dt");
    hip_check(hipEventRecord(c.ev[5], c.stream), "event");
    cfdg::launch_jacobian(c.margs(), c.fargs(), c.nargs, c.bargs(), c.alpha.get(), c.values.get(), c.storage,
                          c.stream);
    hip_check(hipEventRecord(c.ev[6], c.stream), "event");
    c.check_launch();
    c.sync();
    c.have_jacobian = true;
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, c.ev[5], c.ev[6]), "elapsed");
    c.stage_ms[4] = ms;
}

void require_jacobian(const cfdg_ctx& c) {
    if (!c.have_jacobian)
        throw CfdgError(CFDG_ERR_STATE,
                        "no block values in the current storage: cfdg_assemble_jacobian or cfdg_set_bsr_values first");
}

void fetch_dots(cfdg_ctx& c, int n) {
    c.d2h(c.h_dots, c.dots.get(), static_cast<std::size_t>(n));
    c.sync();
}

template <class T>
double dot_device(cfdg_ctx& c, const T* a, const T* b) {
    cfdg::DotPairsT<T> pr{};
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

// Work vectors of one Krylov precision.
template <class T>
struct KVecs {
    T *r, *r0, *p, *v, *s, *t, *ph, *sh;
};
KVecs<double> kvecs64(cfdg_ctx& c) {
    return {c.r.get(), c.r0.get(), c.p.get(), c.v.get(), c.s.get(), c.t.get(), c.ph.get(), c.sh.get()};
}
KVecs<float> kvecs32(cfdg_ctx& c) {
    auto& f = c.f32;
    return {f.r.get(), f.r0.get(), f.p.get(), f.v.get(), f.s.get(), f.t.get(), f.ph.get(), f.sh.get()};
}

template <class T>
void spmv_stored(cfdg_ctx& c, const T* x, T* y, const double* gate = nullptr) {
    cfdg::launch_spmv<T>(c.bargs(), c.values.get(), c.storage, x, y, c.stream, gate);
}

// z = M^-1 r: z = 0, forward groups, backward groups (preconditioners.cpp:171-194).
// The forward sweep reads only the blocks whose column lies in an earlier
// group -- the rows of later groups still hold z = 0 there, so the other
// products add nothing (with the dependency levels these are exactly the
// columns j < i of the reference's sweep) -- and it writes every row, so z
// needs no clearing.  The backward sweep starts with the group the forward
// sweep ended with: those rows read only neighbours that the forward sweep
// already left final, so their backward update equals their forward one bit
// for bit and is skipped.  first_done: the forward sweep's first group was
// already solved by the kernel that formed r (vec_sgs_first).
template <class T>
void precond_apply(cfdg_ctx& c, const Precond& P, const T* r, T* z, const double* gate = nullptr,
                   bool first_done = false) {
    const long long n5 = 5LL * c.nc;
    if (P.kind == CFDG_PRECOND_NONE) {
        hip_check(hipMemcpyAsync(z, r, static_cast<std::size_t>(n5) * sizeof(T), hipMemcpyDeviceToDevice, c.stream),
                  "copy");
        return;
    }
    const bool seq = P.order == CFDG_SGS_SEQUENTIAL;
    const int gf = seq ? 2 : 0;
    const auto b = c.bargs();
    for (int sweep = 0; sweep < 2; ++sweep) {
        const int gi = gf + sweep;
        const auto& off = c.h_gofs[gi];
        const std::uint8_t* lower = sweep == 0 ? c.blk_lower[seq ? 1 : 0].get() : nullptr;
        for (std::size_t k = (sweep == 1 || first_done ? 1 : 0); k + 1 < off.size(); ++k)
            cfdg::launch_sgs_group<T>(c.grows[gi].get() + off[k], off[k + 1] - off[k], b, c.values.get(), c.storage,
                                      c.lu.get(), c.piv.get(), lower, r, z, c.stream, gate);
    }
}

// Diagonal-block LU of the SGS preconditioner (preconditioners.cpp:62-93).
// The singular-pivot flag is read with the solve's result (precond_check):
// a zero pivot makes the solve produce non-finite values, and the error names
// the row either way.
void precond_setup(cfdg_ctx& c, const Precond& P) {
    if (P.kind != CFDG_PRECOND_SGS) return;
    if (P.order == CFDG_SGS_SEQUENTIAL) build_seq_groups(c);
    c.singular.zero(c.stream);
    cfdg::launch_lu_factor(c.nc, c.diag.get(), c.values.get(), c.storage, c.lu.get(), c.lu_f32, c.piv.get(),
                           c.singular.get(), c.stream);
    c.check_launch();
}

void precond_check(cfdg_ctx& c, const Precond& P) {
    if (P.kind != CFDG_PRECOND_SGS) return;
    int sing = 0;
    c.d2h(&sing, c.singular.get(), 1);
    c.sync();
    if (sing != 0)
        throw CfdgError(CFDG_ERR_SINGULAR,
                        "SGS: singular diagonal block at row " + std::to_string(c.new2old[z(sing - 1)]));
}

// Right-preconditioned BiCGSTAB (bicgstab.cpp:42-201) on device vectors of
// type T (FP64: the reference's arithmetic; FP32: the inner solve of the
// defect correction).  x holds the initial guess on entry and the solution on
// exit.  Scalars and dot products are FP64.
template <class T>
cfdg_iter_result bicgstab(cfdg_ctx& c, const cfdg_linparams& lp, const Precond& P, const KVecs<T>& w, T* x,
                          const T* bvec, bool zero_fallback = false, bool x_zero = false) {
    const long long N = 5LL * c.nc;
    const std::size_t bytes = static_cast<std::size_t>(N) * sizeof(T);
    cfdg_iter_result res{CFDG_MAX_ITERATIONS, 0, 0.0, 0.0};
    // two state blocks: an iteration that ends in launch_xr_next reads one and
    // leaves the next iteration's state in the other
    double* const kb[2] = {c.ks.get(), c.ks.get() + cfdg::kKsCount};
    int cur = 0;
    double* ks = kb[0];
    double* h = c.h_ks;
    // r = b - A x (r = b for a zero guess), norms, tolerance and status on the
    // device: no host round trip before the first batch of iterations
    {
        cfdg::DotPairsT<T> pr{};
        pr.a[0] = bvec;
        pr.b[0] = bvec;
        if (x_zero) {
            hip_check(hipMemcpyAsync(w.r, bvec, bytes, hipMemcpyDeviceToDevice, c.stream), "copy");
        } else {
            spmv_stored<T>(c, x, w.v);
            cfdg::launch_axpy_copy<T>(N, bvec, -1.0, w.v, w.r, c.stream);
            pr.a[1] = w.r;
            pr.b[1] = w.r;
        }
        cfdg::launch_dots(N, x_zero ? 1 : 2, pr, c.partial.get(), c.dots.get(), c.stream);
        cfdg::launch_bicg_init(ks, c.dots.get(), x_zero ? 1 : 0, zero_fallback ? 1 : 0, lp.abs_tol, lp.rel_tol,
                               lp.max_iter, c.stream);
        if (zero_fallback) cfdg::launch_bicg_reject<T>(N, ks, bvec, x, w.r, c.stream);
    }
    int it = 0;
    double rnorm = 0.0;
    {
        hip_check(hipMemcpyAsync(w.r0, w.r, bytes, hipMemcpyDeviceToDevice, c.stream), "copy");
        hip_check(hipMemsetAsync(w.v, 0, bytes, c.stream), "memset");
        // The recurrence runs on the device: scalars in the state block ks,
        // iterations enqueued in batches and the state read back once per
        // batch; iterations past convergence or breakdown are no-ops.
        const int slot = (sizeof(T) == sizeof(double)) ? 0 : 1;
        int batch = std::max(1, c.last_iters[slot] - 1);
        const bool sgs = P.kind == CFDG_PRECOND_SGS;
        const std::uint8_t* first = c.row_first[P.order == CFDG_SGS_SEQUENTIAL ? 1 : 0].get();
        long long enq = 0;
        double* rr = c.dots.get() + 5;
        // with SGS, every iteration but the last of a batch (and the ones with
        // the true residual norm) ends with launch_xr_next, which also forms
        // the next iteration's p and the first group of its M^-1 p
        bool p_ready = false;
        while (true) {
            for (int b = 0; b < batch && enq <= lp.max_iter; ++b) {
                ++enq;
                ks = kb[cur];
                // p, and with SGS the first forward group of p^ = M^-1 p, in one launch
                if (p_ready) {
                    // formed by the previous iteration's launch_xr_next
                } else if (sgs) {
                    cfdg::launch_vec_sgs_first<T>(0, c.nc, ks, w.r, w.v, w.p, first, c.storage, c.lu.get(),
                                                  c.piv.get(), w.ph, c.stream);
                } else {
                    cfdg::launch_update_p_g<T>(N, ks, w.r, w.v, w.p, c.stream);
                }
                precond_apply<T>(c, P, w.p, w.ph, ks, sgs);
                spmv_stored<T>(c, w.ph, w.v, ks);
                {
                    cfdg::DotPairsT<T> pr{};
                    pr.a[0] = w.r0;
                    pr.b[0] = w.v;
                    cfdg::launch_dots_step(N, 1, pr, c.partial.get(), c.dots.get(), ks, cfdg::kStepAlpha, c.stream);
                }
                if (sgs) {
                    cfdg::launch_vec_sgs_first<T>(1, c.nc, ks, w.r, w.v, w.s, first, c.storage, c.lu.get(),
                                                  c.piv.get(), w.sh, c.stream);
                } else {
                    cfdg::launch_axpy_alpha_g<T>(N, ks, w.r, w.v, w.s, c.stream);
                }
                precond_apply<T>(c, P, w.s, w.sh, ks, sgs);
                spmv_stored<T>(c, w.sh, w.t, ks);
                {
                    cfdg::DotPairsT<T> pr{};
                    const T* A[5] = {w.t, w.t, w.r0, w.r0, w.s};
                    const T* B[5] = {w.t, w.s, w.t, w.s, w.s};
                    for (int j = 0; j < 5; ++j) {
                        pr.a[j] = A[j];
                        pr.b[j] = B[j];
                    }
                    cfdg::launch_dots_step(N, 5, pr, c.partial.get(), c.dots.get(), ks, cfdg::kStepOmega, c.stream);
                }
                const bool next_in_batch = b + 1 < batch && enq <= lp.max_iter;
                if (sgs && next_in_batch && enq % 32 != 0) {
                    cfdg::launch_xr_next<T>(c.nc, ks, kb[1 - cur], w.ph, w.sh, w.s, w.t, x, w.r, w.v, w.p, first,
                                            c.storage, c.lu.get(), c.piv.get(), c.stream);
                    cur = 1 - cur;
                    p_ready = true;
                    continue;
                }
                p_ready = false;
                cfdg::launch_update_xr_g<T>(N, ks, w.ph, w.sh, w.s, w.t, x, w.r, c.stream);
                if (enq % 32 == 0) {  // true residual norm every 32 iterations
                    cfdg::DotPairsT<T> pr{};
                    pr.a[0] = w.r;
                    pr.b[0] = w.r;
                    cfdg::launch_dots_step(N, 1, pr, c.partial.get(), rr, ks, cfdg::kStepPost, c.stream);
                } else {
                    cfdg::launch_bicg_post(ks, c.stream);
                }
            }
            c.check_launch();
            ks = kb[cur];
            c.d2h(h, ks, cfdg::kKsCount);
            c.sync();
            if (h[cfdg::kKsStatus] != 0.0 || enq > lp.max_iter) break;
            batch = 2;
        }
        it = static_cast<int>(h[cfdg::kKsIt]);
        rnorm = h[cfdg::kKsRnorm];
        const double st = h[cfdg::kKsStatus];
        res.status = (st == cfdg::kKsConverged) ? CFDG_CONVERGED
                                                 : (st == cfdg::kKsDiverged ? CFDG_DIVERGED : CFDG_MAX_ITERATIONS);
        res.initial_residual = h[cfdg::kKsRnorm0];
        if (zero_fallback) {
            if (h[cfdg::kKsReject] != 0.0)
                ++c.warm_rejected;
            else
                ++c.warm_used;
        }
        if (it > 0) c.last_iters[slot] = it;
    }
    c.check_launch();
    c.sync();
    res.iterations = it;
    res.final_residual = rnorm;
    return res;
}

// FP64 defect correction around FP32 BiCGSTAB + SGS (the FP32 precision form):
// x (FP64, initial guess on entry), b (FP64).  Every pass forms r = b - A x in
// FP64 arithmetic over the stored blocks and solves A e = r in FP32 from e = 0.
cfdg_iter_result defect_correction(cfdg_ctx& c, const cfdg_linparams& lp, const Precond& P, double* x,
                                   const double* bvec, bool zero_fallback = false, bool x_zero = false) {
    const long long N = 5LL * c.nc;
    cfdg_iter_result res{CFDG_MAX_ITERATIONS, 0, 0.0, 0.0};
    const auto outer_residual = [&]() {
        spmv_stored<double>(c, x, c.v.get());
        cfdg::launch_axpy_copy<double>(N, bvec, -1.0, c.v.get(), c.r.get(), c.stream);
        return std::sqrt(dot_device<double>(c, c.r.get(), c.r.get()));
    };
    // ||b|| and the first outer residual in one reduction (r = b for a zero guess)
    double bnorm = 0.0, rnorm = 0.0;
    {
        cfdg::DotPairsT<double> pr{};
        pr.a[0] = bvec;
        pr.b[0] = bvec;
        if (x_zero) {
            hip_check(hipMemcpyAsync(c.r.get(), bvec, static_cast<std::size_t>(N) * sizeof(double),
                                     hipMemcpyDeviceToDevice, c.stream), "copy");
        } else {
            spmv_stored<double>(c, x, c.v.get());
            cfdg::launch_axpy_copy<double>(N, bvec, -1.0, c.v.get(), c.r.get(), c.stream);
            pr.a[1] = c.r.get();
            pr.b[1] = c.r.get();
        }
        cfdg::launch_dots(N, x_zero ? 1 : 2, pr, c.partial.get(), c.dots.get(), c.stream);
        fetch_dots(c, x_zero ? 1 : 2);
        bnorm = std::sqrt(c.h_dots[0]);
        rnorm = x_zero ? bnorm : std::sqrt(c.h_dots[1]);
    }
    if (zero_fallback) {
        if (!(rnorm <= bnorm)) {
            c.x.zero(c.stream);
            hip_check(hipMemcpyAsync(c.r.get(), bvec, static_cast<std::size_t>(N) * sizeof(double),
                                     hipMemcpyDeviceToDevice, c.stream), "copy");
            rnorm = bnorm;
            ++c.warm_rejected;
        } else {
            ++c.warm_used;
        }
    }
    const double tol = std::max(lp.abs_tol, lp.rel_tol * ((bnorm > 0.0) ? bnorm : 1.0));
    res.initial_residual = rnorm;
    const KVecs<float> w = kvecs32(c);
    int total = 0;
    for (int pass = 0;; ++pass) {
        if (!std::isfinite(rnorm)) {
            res.status = CFDG_DIVERGED;
            break;
        }
        if (rnorm <= tol) {
            res.status = CFDG_CONVERGED;
            break;
        }
        if (pass >= c.prec.max_outer || total >= lp.max_iter) {
            res.status = CFDG_MAX_ITERATIONS;
            break;
        }
        cfdg::launch_cast_f64_to_f32(N, c.r.get(), c.f32.b.get(), c.stream);
        c.f32.x.zero(c.stream);
        cfdg_linparams in = lp;
        // stop the inner solve at max(inner_rel_tol ||r||, 0.9 tol): no further
        // than the caller's tolerance needs (the FP64 outer residual of the
        // update equals the inner residual up to FP32 rounding), and at least a
        // factor inner_rel_tol per pass when the tolerance lies beyond FP32 reach
        in.rel_tol = c.prec.inner_rel_tol;
        in.abs_tol = 0.9 * tol;
        in.max_iter = lp.max_iter - total;
        const cfdg_iter_result ir = bicgstab<float>(c, in, P, w, c.f32.x.get(), c.f32.b.get(), false, true);
        total += ir.iterations;
        if (ir.iterations == 0 && ir.status != CFDG_CONVERGED) {
            res.status = CFDG_DIVERGED;
            break;
        }
        cfdg::launch_add_f32(N, c.f32.x.get(), x, c.stream);
        rnorm = outer_residual();
    }
    c.check_launch();
    c.sync();
    res.iterations = total;
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
    require_jacobian(c);
    const std::size_t n5 = 5ULL * static_cast<std::size_t>(c.nc);
    if (hb) {
        const auto d = c.vec_to_device(hb);
        c.h2d(c.b.get(), d.data(), n5);
        c.sync();
    } else {
        if (!c.have_residual) throw CfdgError(CFDG_ERR_STATE, "implicit right-hand side needs a residual");
        cfdg::launch_rhs(c.nc, c.nc, c.res.get(), c.b.get(), c.stream);
    }
    bool learned = false, x_zero = false;
    if (hx0) {
        const auto d = c.vec_to_device(hx0);
        c.h2d(c.x.get(), d.data(), n5);
        c.sync();
    } else if (!hb && c.warm) {
        learned = cfdg::gnn_warm_guess(c, c.x.get());
    } else {
        c.x.zero(c.stream);
        x_zero = true;
    }
    Precond P;
    P.kind = lp.preconditioner;
    P.order = lp.sgs_order;
    hip_check(hipEventRecord(c.ev[5], c.stream), "event");
    precond_setup(c, P);
    const cfdg_iter_result r = (c.prec.krylov == CFDG_PREC_FP32)
                                   ? defect_correction(c, lp, P, c.x.get(), c.b.get(), learned, x_zero)
                                   : bicgstab<double>(c, lp, P, kvecs64(c), c.x.get(), c.b.get(), learned, x_zero);
    hip_check(hipEventRecord(c.ev[6], c.stream), "event");
    c.sync();
    precond_check(c, P);
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

void setup_context(cfdg_ctx& c, const cfdg::Mesh& m, const cfdg::MeshConnectivity* cn, const cfdg_numerics* num,
                   const cfdg_bc* bcs, int32_t n_bcs) {
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
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&c.h_ks), cfdg::kKsCount * sizeof(double)), "hipHostMalloc");

    c.num = *num;
    c.nc = m.n_cells;
    c.nf = m.n_faces;
    c.ninner = m.n_inner_faces;
    c.nbf = m.n_faces - m.n_inner_faces;
    c.ntot = c.nc + c.nbf;
    c.npatch = n_bcs;
    c.nargs.gas = cfdg::make_gas(num->gamma, num->gas_constant, num->prandtl, num->prandtl_turb);
    c.nargs.viscous = num->flow_model == CFDG_VISCOUS_FLOW;
    c.nargs.muscl = num->reconstruction == CFDG_MUSCL;
    c.nargs.limiter = num->limiter;
    c.nargs.gradient = num->gradient;
    c.nargs.k3 = num->venkat_k * num->venkat_k * num->venkat_k;
    for (int32_t p = 0; p < n_bcs; ++p) c.bcs.bc[p] = resolve_bc(bcs[p], num->gamma, num->gas_constant);
    hipStream_t s = c.stream;

    // mesh-numbered connectivity (built with the mesh when it holds one),
    // device tiling and permutation
    if (cn && (cn->cell_faces.off.size() != z(c.nc) + 1 || cn->bsr.diag.size() != z(c.nc)))
        throw CfdgError(CFDG_ERR_STATE, "the mesh's connectivity does not match its cells");
    cfdg::Csr cf_own;
    if (!cn) cf_own = cfdg::cell_faces(m);
    const cfdg::Csr& cf_mesh = cn ? cn->cell_faces : cf_own;
    c.ccf_mesh = cn ? cn->cell_cells_face : cfdg::cell_cells_by_face(m);
    const cfdg::Tiling tl = cfdg::tile_cells(m, c.ccf_mesh, cf_mesh, cfdg::kTileCells, cfdg::kTileFaces);
    c.new2old = tl.new2old;
    c.old2new = tl.old2new;
    c.ntiles = static_cast<int>(tl.tile_off.size()) - 1;

    // face planes (mesh face order) and (owner, neighbour or ghost) in device numbering
    std::vector<i32> owner_dev(z(c.nf)), other_dev(z(c.nf));
    std::vector<double> fph;
    std::vector<int2> fcells(z(c.nf));
    c.fld = (static_cast<long long>(c.nf) + 63) / 64 * 64;
    {
        std::vector<double> dx, dy, dz, di;
        cfdg::face_distances(m, dx, dy, dz, di);
        fph.assign(z(c.fld) * cfdg::kFacePlanes, 0.0);
        const auto put = [&](int k, int f, double v) { fph[z(cfdg::plane_at(k, c.fld, f))] = v; };
        for (int f = 0; f < c.nf; ++f) {
            put(cfdg::kFN, f, m.face_nx[z(f)]);
            put(cfdg::kFN + 1, f, m.face_ny[z(f)]);
            put(cfdg::kFN + 2, f, m.face_nz[z(f)]);
            put(cfdg::kFA, f, m.face_area[z(f)]);
            put(cfdg::kFC, f, m.face_cx[z(f)]);
            put(cfdg::kFC + 1, f, m.face_cy[z(f)]);
            put(cfdg::kFC + 2, f, m.face_cz[z(f)]);
            put(cfdg::kFD, f, dx[z(f)]);
            put(cfdg::kFD + 1, f, dy[z(f)]);
            put(cfdg::kFD + 2, f, dz[z(f)]);
            put(cfdg::kFDinv, f, di[z(f)]);
            owner_dev[z(f)] = c.old2new[z(m.face_owner[z(f)])];
            const i32 other = (f < c.ninner) ? c.old2new[z(m.face_neigh[z(f)])] : c.nc + (f - c.ninner);
            other_dev[z(f)] = other;
            fcells[z(f)] = int2{owner_dev[z(f)], other};
        }
        c.fp.upload(fph, s);
        c.fcells.upload(fcells, s);
        c.sync();
    }
    {
        const cfdg::Csr cf_dev = to_device_rows(cf_mesh, c.new2old, nullptr);
        const cfdg::TileFaces tf = cfdg::tile_faces(tl, cf_dev, owner_dev, other_dev);
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
        // the flux kernel's copy of the face planes, in the order of the tiles'
        // face lists: normal, area, centroid, and the packed (owner, other) ids
        const long long nte = static_cast<long long>(tf.faces.size());
        c.tld = (nte + 63) / 64 * 64;
        std::vector<double> tph(z(c.tld) * cfdg::kTilePlanes, 0.0);
        for (long long e = 0; e < nte; ++e) {
            const i32 f = tf.faces[z(e)];
            for (int k = 0; k < cfdg::kTCells; ++k)
                tph[z(cfdg::plane_at(k, c.tld, e))] = fph[z(cfdg::plane_at(k, c.fld, f))];
            const std::uint64_t ids = static_cast<std::uint32_t>(fcells[z(f)].x) |
                                      (static_cast<std::uint64_t>(static_cast<std::uint32_t>(fcells[z(f)].y)) << 32);
            std::memcpy(&tph[z(cfdg::plane_at(cfdg::kTCells, c.tld, e))], &ids, sizeof ids);
        }
        c.tf.upload(tph, s);
        // entry of every face in the face list of the tile that holds its owner
        std::vector<i32> fentry(z(c.nf), -1);
        for (int k = 0; k < c.ntiles; ++k)
            for (i32 e = tf.tile_face_off[z(k)]; e < tf.tile_face_off[z(k) + 1]; ++e) {
                const i32 f = tf.faces[z(e)];
                const i32 o = owner_dev[z(f)];
                if (o >= tl.tile_off[z(k)] && o < tl.tile_off[z(k) + 1]) fentry[z(f)] = e;
            }
        c.face_entry.upload(fentry, s);
        const cfdg::Csr ccf_dev = to_device_rows(c.ccf_mesh, c.new2old, &c.old2new);
        c.ccf_off.upload(ccf_dev.off, s);
        c.ccf.upload(ccf_dev.idx, s);
        std::vector<int> bfb(z(c.nbf));
        for (int i = 0; i < c.nbf; ++i) bfb[z(i)] = m.face_patch[z(c.ninner + i)];
        c.bface_bc.upload(bfb, s);
        c.sync();
    }
    if (num->gradient == CFDG_LEAST_SQUARES_NODE) {
        cfdg::Csr st_own;
        if (!(cn && cn->has_node_stencil)) st_own = cfdg::cell_cells_by_node(m);
        const cfdg::Csr st = to_device_rows(cn && cn->has_node_stencil ? cn->cell_cells_node : st_own, c.new2old,
                                            &c.old2new);
        c.ccn_off.upload(st.off, s);
        c.ccn.upload(st.idx, s);
        // LDS-staged stencils: the flux tiles when their cells + halo fit the
        // kernel's slots (the wide form when the narrow one does not hold
        // them), else consecutive sub-ranges of them that fit the wide form
        const cfdg::Csr ccf_dev = to_device_rows(c.ccf_mesh, c.new2old, &c.old2new);
        std::vector<i32> goff = tl.tile_off;
        if (cfdg::tile_max_slots(goff, st, ccf_dev) > cfdg::kGradSlotsBig)
            goff = cfdg::split_tiles(tl.tile_off, st, ccf_dev, cfdg::kGradSlotsBig);
        const cfdg::TileStencil ts = cfdg::tile_stencil(goff, st, ccf_dev);
        c.gslots = (ts.max_slots <= cfdg::kGradSlots) ? cfdg::kGradSlots : cfdg::kGradSlotsBig;
        if (ts.max_slots <= cfdg::kGradSlotsBig) {
            int most = 0;
            for (std::size_t k = 0; k + 1 < goff.size(); ++k) most = std::max(most, goff[k + 1] - goff[k]);
            c.ngtiles = static_cast<int>(goff.size()) - 1;
            c.gblock = std::min(cfdg::kTileCells, (most + 63) / 64 * 64);
            c.gtile_off.upload(goff, s);
            c.tile_halo_off.upload(ts.halo_off, s);
            c.tile_halo.upload(ts.halo, s);
            c.ccn_slot.upload(ts.ccn_slot, s);
            c.ccf_slot.upload(ts.ccf_slot, s);
            c.ccn_base.upload(ts.ccn_base, s);
            c.ccf_base.upload(ts.ccf_base, s);
            // limiter face vectors x_f - x_c of every cell face (ascending face
            // index), per gradient tile transposed like the slot lists: tile t
            // holds 3 components x W_t faces x nt cells, W_t = its longest list
            const cfdg::Csr cf_dev = to_device_rows(cf_mesh, c.new2old, nullptr);
            std::vector<long long> fvb(goff.size(), 0);
            for (std::size_t k = 0; k + 1 < goff.size(); ++k) {
                i32 w = 0;
                for (i32 cc = goff[k]; cc < goff[k + 1]; ++cc) w = std::max(w, cf_dev.off[z(cc) + 1] - cf_dev.off[z(cc)]);
                fvb[k + 1] = fvb[k] + 3LL * w * (goff[k + 1] - goff[k]);
            }
            std::vector<double> fv(z(fvb.back()), 0.0);
            for (std::size_t k = 0; k + 1 < goff.size(); ++k) {
                const i32 nt = goff[k + 1] - goff[k];
                const long long w = (fvb[k + 1] - fvb[k]) / (3LL * nt);
                for (i32 cc = goff[k]; cc < goff[k + 1]; ++cc) {
                    const std::size_t mc = z(c.new2old[z(cc)]);
                    for (i32 e = cf_dev.off[z(cc)]; e < cf_dev.off[z(cc) + 1]; ++e) {
                        const std::size_t f = z(cf_dev.idx[z(e)]);
                        const long long at = fvb[k] + (e - cf_dev.off[z(cc)]) * static_cast<long long>(nt) + (cc - goff[k]);
                        fv[z(at)] = m.face_cx[f] - m.cell_cx[mc];
                        fv[z(at + w * nt)] = m.face_cy[f] - m.cell_cy[mc];
                        fv[z(at + 2 * w * nt)] = m.face_cz[f] - m.cell_cz[mc];
                    }
                }
            }
            c.face_vec.upload(fv, s);
            c.face_vec_base.upload(fvb, s);
            c.sync();  // fv and fvb are released at the end of this block
        }
        c.sync();
    }
    // cell planes: static centroid and volume; everything else zero
    {
        c.ld = (static_cast<long long>(c.ntot) + 63) / 64 * 64;
        std::vector<double> cp(z(c.ld) * cfdg::kCellPlanes, 0.0);
        const auto put = [&](int k, int r, double v) { cp[z(cfdg::plane_at(k, c.ld, r))] = v; };
        for (int r = 0; r < c.nc; ++r) {
            const std::size_t o = z(c.new2old[z(r)]);
            put(cfdg::kRC, r, m.cell_cx[o]);
            put(cfdg::kRC + 1, r, m.cell_cy[o]);
            put(cfdg::kRC + 2, r, m.cell_cz[o]);
            put(cfdg::kRVol, r, m.cell_vol[o]);
        }
        c.cs.upload(cp, s);
        // inverse least-squares normal matrices: geometry only, once per mesh
        if (num->gradient == CFDG_LEAST_SQUARES_NODE) {
            cfdg::launch_lsq_setup(c.margs(), c.fargs(), s);
            c.check_launch();
        }
        c.sync();
    }
    // BSR: the mesh-order pattern (bsr_helpers.hpp:16-59) and its device rows
    {
        cfdg::BsrPattern bp_own;
        if (!cn) bp_own = cfdg::bsr_pattern(c.ccf_mesh, c.nc);
        const cfdg::BsrPattern& bp = cn ? cn->bsr : bp_own;
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
        // off-diagonal block of every interior cell face, in the device cell ->
        // face lists (the Jacobian kernel's store positions)
        {
            const cfdg::Csr cf_dev = to_device_rows(cf_mesh, c.new2old, nullptr);
            std::vector<i32> blk(cf_dev.idx.size(), -1);
            for (int r = 0; r < c.nc; ++r)
                for (i32 e = cf_dev.off[z(r)]; e < cf_dev.off[z(r) + 1]; ++e) {
                    const i32 f = cf_dev.idx[z(e)];
                    if (f >= c.ninner) continue;
                    const i32 o = owner_dev[z(f)] == r ? other_dev[z(f)] : owner_dev[z(f)];
                    i32 k = rp[z(r)];
                    while (k < rp[z(r) + 1] && cols[z(k)] != o) ++k;
                    if (k == rp[z(r) + 1]) throw CfdgError(CFDG_ERR_MESH, "face neighbour missing from the BSR row");
                    blk[z(e)] = k;
                }
            c.cf_blk.upload(blk, s);
            c.sync();  // blk is released at the end of this block
        }
        i32 ncol = 0;
        const auto col = cfdg::greedy_colours(c.ccf_mesh, c.nc, ncol);
        upload_groups(c, 0, col, ncol, false);
        upload_groups(c, 1, col, ncol, true);
        upload_lower_flags(c, 0, col);
        c.sync();
    }
    const std::size_t ncs = z(c.nc);
    c.U.alloc(5 * ncs);
    c.res.alloc(5 * ncs);
    c.lam.alloc(ncs);
    c.bflux.alloc(5 * static_cast<std::size_t>(std::max(c.nbf, 1)));
    c.bint.alloc(5 * static_cast<std::size_t>(std::max(c.npatch, 1)));
    c.dt.alloc(ncs);
    c.alpha.alloc(ncs);
    c.U.zero(s);
    c.res.zero(s);
    c.lam.zero(s);
    c.values.alloc(c.h_cols.size() * 25 * sizeof(double));
    c.values.zero(s);
    c.lu.alloc(ncs * 25 * sizeof(double));
    c.piv.alloc(ncs * 5);
    c.singular.alloc(1);
    for (DBuf<double>* v : {&c.b, &c.x, &c.r, &c.r0, &c.p, &c.v, &c.s, &c.t, &c.ph, &c.sh}) {
        v->alloc(5 * ncs);
        v->zero(s);
    }
    c.partial.alloc(static_cast<std::size_t>(cfdg::kReduceBlocks) * 5);
    c.dots.alloc(8);
    c.ks.alloc(2 * cfdg::kKsCount);
    c.sync();
}

// BackwardEuler::advance (implicit_euler.hpp:34-76): residual, dt, Jacobian,
// (alpha I + J) du = -R, U <- U + du; the step proceeds whatever the linear
// solver's status (implicit_euler.hpp:45-47).  out->l2 = norms of R(U^n).
void implicit_step(cfdg_ctx& c, const cfdg_linparams& lp, cfdg_step_result* out) {
    eval_residual(c);
    compute_dt(c);
    cfdg::gnn_step_features(c);
    assemble_jacobian(c);
    cfdg_iter_result lr{};
    linear_solve(c, lp, nullptr, nullptr, nullptr, &lr);
    double l2[5];
    residual_norms(c, l2);
    cfdg::launch_state_update(c.nc, c.nc, c.x.get(), c.U.get(), c.stream);
    // module block: copied (implicit_euler.hpp:64-72) or point-implicit, then clamped
    if (c.sa_on())
        cfdg::launch_sa_update(c.nc, c.turb.update == CFDG_TURB_POINT_IMPLICIT, c.alpha.get(), c.sargs(), c.stream);
    c.check_launch();
    c.sync();
    c.have_residual = false;
    c.have_dt = false;
    c.have_prev_du = true;
    for (int v = 0; v < 5; ++v) out->l2[v] = l2[v];
    out->linear = lr;
}

std::size_t storage_bytes(int storage) {
    return storage == cfdg::kStoreF64 ? sizeof(double) : (storage == cfdg::kStoreF32 ? sizeof(float) : sizeof(cfdg::bf16));
}

// Stored values per block (bfloat16 blocks are padded, see BlockLayout).
std::size_t block_stride(int storage) { return storage == cfdg::kStoreBF16 ? cfdg::BlockLayout<cfdg::bf16>::block : 25; }

// Device-order block values as double (25 per block), whatever the storage.
std::vector<double> download_values(cfdg_ctx& c) {
    const std::size_t n = c.blk_dev.size() * 25;
    std::vector<double> out(n);
    if (c.storage == cfdg::kStoreF64) {
        c.d2h(out.data(), reinterpret_cast<const double*>(c.values.get()), n);
        c.sync();
    } else if (c.storage == cfdg::kStoreF32) {
        std::vector<float> h(n);
        c.d2h(h.data(), reinterpret_cast<const float*>(c.values.get()), n);
        c.sync();
        for (std::size_t i = 0; i < n; ++i) out[i] = cfdg::to_f64(h[i]);
    } else {
        constexpr int kB = cfdg::BlockLayout<cfdg::bf16>::block, kR = cfdg::BlockLayout<cfdg::bf16>::row;
        std::vector<cfdg::bf16> h(c.blk_dev.size() * kB);
        c.d2h(h.data(), reinterpret_cast<const cfdg::bf16*>(c.values.get()), h.size());
        c.sync();
        for (std::size_t k = 0; k < c.blk_dev.size(); ++k)
            for (int i = 0; i < 25; ++i) out[k * 25 + z(i)] = cfdg::to_f64(h[k * kB + z((i / 5) * kR + i % 5)]);
    }
    return out;
}

void upload_values(cfdg_ctx& c, const std::vector<double>& h) {
    const std::size_t n = h.size();
    if (c.storage == cfdg::kStoreF64) {
        c.h2d(reinterpret_cast<double*>(c.values.get()), h.data(), n);
        c.sync();
    } else if (c.storage == cfdg::kStoreF32) {
        std::vector<float> t(n);
        for (std::size_t i = 0; i < n; ++i) t[i] = cfdg::cvt<float>(h[i]);
        c.h2d(reinterpret_cast<float*>(c.values.get()), t.data(), n);
        c.sync();
    } else {
        constexpr int kB = cfdg::BlockLayout<cfdg::bf16>::block, kR = cfdg::BlockLayout<cfdg::bf16>::row;
        std::vector<cfdg::bf16> t(n / 25 * kB, cfdg::bf16{0});
        for (std::size_t k = 0; k < n / 25; ++k)
            for (int i = 0; i < 25; ++i) t[k * kB + z((i / 5) * kR + i % 5)] = cfdg::cvt<cfdg::bf16>(h[k * 25 + z(i)]);
        c.h2d(reinterpret_cast<cfdg::bf16*>(c.values.get()), t.data(), t.size());
        c.sync();
    }
}

void set_precision(cfdg_ctx& c, const cfdg_precision& p) {
    const bool st_ok = p.block_storage == CFDG_PREC_FP64 || p.block_storage == CFDG_PREC_FP32 ||
                       p.block_storage == CFDG_PREC_BF16;
    if (!st_ok) throw CfdgError(CFDG_ERR_CONFIG, "unknown block storage");
    if (p.krylov != CFDG_PREC_FP64 && p.krylov != CFDG_PREC_FP32) throw CfdgError(CFDG_ERR_CONFIG, "unknown Krylov precision");
    if (p.krylov == CFDG_PREC_FP64 && p.block_storage != CFDG_PREC_FP64)
        throw CfdgError(CFDG_ERR_CONFIG, "FP64 Krylov needs FP64 block storage");
    if (p.krylov == CFDG_PREC_FP32 && (p.max_outer < 1 || !(p.inner_rel_tol > 0.0 && p.inner_rel_tol < 1.0)))
        throw CfdgError(CFDG_ERR_CONFIG, "FP32 Krylov needs max_outer >= 1 and 0 < inner_rel_tol < 1");
    const int storage = p.block_storage;  // CFDG_PREC_* == kStore*
    const std::size_t nnzb = c.blk_dev.size();
    const std::size_t ncs = z(c.nc);
    if (storage != c.storage || c.values.size() != nnzb * block_stride(storage) * storage_bytes(storage)) {
        c.values.alloc(0);
        c.values.alloc(nnzb * block_stride(storage) * storage_bytes(storage));
        c.values.zero(c.stream);
        c.have_jacobian = false;
    }
    const int lu_f32 = (p.krylov == CFDG_PREC_FP32) ? 1 : 0;
    c.lu.alloc(0);
    c.lu.alloc(ncs * 25 * (lu_f32 ? sizeof(float) : sizeof(double)));
    if (p.krylov == CFDG_PREC_FP32) {
        auto& f = c.f32;
        for (DBuf<float>* v : {&f.b, &f.x, &f.r, &f.r0, &f.p, &f.v, &f.s, &f.t, &f.ph, &f.sh}) {
            if (v->size() != 5 * ncs) v->alloc(5 * ncs);
            v->zero(c.stream);
        }
    } else {
        auto& f = c.f32;
        for (DBuf<float>* v : {&f.b, &f.x, &f.r, &f.r0, &f.p, &f.v, &f.s, &f.t, &f.ph, &f.sh}) v->alloc(0);
    }
    c.sync();
    c.storage = storage;
    c.lu_f32 = lu_f32;
    c.prec = p;
    c.prec.reserved = 0;
}

}  // namespace

extern "C" {

int cfdg_create(const cfdg_mesh* mesh, const cfdg_numerics* num, const cfdg_bc* bcs, int32_t n_bcs, cfdg_ctx** out) {
    if (!mesh || !num || !out || (n_bcs > 0 && !bcs)) return CFDG_ERR_ARG;
    *out = nullptr;
    auto* c = new cfdg_ctx();
    const int rc = guarded(c, [&] { setup_context(*c, mesh->m, mesh->conn.get(), num, bcs, n_bcs); });
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
        c->have_prev_du = false;
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
        c->have_prev_du = false;
        // the uniform state is the freestream of the learned warm start's normalisation
        c->ref_rho = rho;
        c->ref_p = p;
        for (int k = 0; k < 3; ++k) c->ref_vel[k] = vel[k];
        c->have_ref = true;
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
        const auto h = c->download_planes();
        for (int r = 0; r < c->ntot; ++r) {
            const long long o = (r < c->nc) ? c->new2old[z(r)] : r;
            for (int v = 0; v < 5; ++v) q[v * ld + o] = h[z(cfdg::plane_at(cfdg::kRQ + v, c->ld, r))];
        }
    });
}

int cfdg_get_gradients(cfdg_ctx* c, double* grad, int64_t ld, double* phi, int64_t ld_phi) {
    return guarded(c, [&] {
        if (grad && ld < c->ntot) throw CfdgError(CFDG_ERR_ARG, "get_gradients: ld < n_total");
        if (phi && ld_phi < c->nc) throw CfdgError(CFDG_ERR_ARG, "get_gradients: ld_phi < n_cells");
        const auto h = c->download_planes();
        const auto at = [&](int k, int r) { return h[z(cfdg::plane_at(k, c->ld, r))]; };
        for (int r = 0; r < c->ntot; ++r) {
            const long long o = (r < c->nc) ? c->new2old[z(r)] : r;
            if (grad)
                for (int k = 0; k < 15; ++k) grad[k * ld + o] = at(cfdg::kRG + k, r);
            if (phi && r < c->nc)
                for (int v = 0; v < 5; ++v) phi[v * ld_phi + o] = at(cfdg::kRPhi + v, r);
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
        const std::vector<double> h = download_values(*c);
        for (std::size_t k = 0; k < c->blk_dev.size(); ++k)
            std::memcpy(values + k * 25, h.data() + z(c->blk_dev[k]) * 25, 25 * sizeof(double));
    });
}

int cfdg_set_bsr_values(cfdg_ctx* c, const double* values) {
    return guarded(c, [&] {
        if (!values) throw CfdgError(CFDG_ERR_ARG, "null input");
        std::vector<double> h(c->blk_dev.size() * 25);
        for (std::size_t k = 0; k < c->blk_dev.size(); ++k)
            std::memcpy(h.data() + z(c->blk_dev[k]) * 25, values + k * 25, 25 * sizeof(double));
        upload_values(*c, h);
        c->have_jacobian = true;
    });
}

int cfdg_set_precision(cfdg_ctx* c, const cfdg_precision* prec) {
    return guarded(c, [&] {
        if (!prec) throw CfdgError(CFDG_ERR_ARG, "null precision");
        set_precision(*c, *prec);
    });
}

int cfdg_get_precision(cfdg_ctx* c, cfdg_precision* prec) {
    return guarded(c, [&] {
        if (!prec) throw CfdgError(CFDG_ERR_ARG, "null output");
        *prec = c->prec;
    });
}

int cfdg_linear_solve(cfdg_ctx* c, const cfdg_linparams* lp, const double* b, const double* x0, double* x,
                      cfdg_iter_result* res) {
    return guarded(c, [&] {
        if (!lp) throw CfdgError(CFDG_ERR_ARG, "null linear parameters");
        linear_solve(*c, *lp, b, x0, x, res);
        c->have_prev_du = false;
    });
}

int cfdg_sgs_apply(cfdg_ctx* c, int32_t sgs_order, const double* r, double* zout) {
    return guarded(c, [&] {
        if (!r || !zout) throw CfdgError(CFDG_ERR_ARG, "null vector");
        if (sgs_order != CFDG_SGS_MULTICOLOUR && sgs_order != CFDG_SGS_SEQUENTIAL)
            throw CfdgError(CFDG_ERR_CONFIG, "unknown SGS order");
        require_jacobian(*c);
        const std::size_t n5 = 5ULL * z(c->nc);
        Precond P;
        P.kind = CFDG_PRECOND_SGS;
        P.order = sgs_order;
        precond_setup(*c, P);
        const auto d = c->vec_to_device(r);
        c->h2d(c->s.get(), d.data(), n5);
        if (c->prec.krylov == CFDG_PREC_FP32) {
            cfdg::launch_cast_f64_to_f32(static_cast<long long>(n5), c->s.get(), c->f32.s.get(), c->stream);
            precond_apply<float>(*c, P, c->f32.s.get(), c->f32.sh.get());
            cfdg::launch_cast_f32_to_f64(static_cast<long long>(n5), c->f32.sh.get(), c->sh.get(), c->stream);
        } else {
            precond_apply<double>(*c, P, c->s.get(), c->sh.get());
        }
        std::vector<double> h(n5);
        c->d2h(h.data(), c->sh.get(), n5);
        c->check_launch();
        c->sync();
        precond_check(*c, P);
        c->vec_to_mesh(h, zout);
    });
}

int cfdg_sgs_row_groups(cfdg_ctx* c, int32_t sgs_order, int32_t* fwd_group, int32_t* bwd_group,
                        int32_t* n_groups) {
    return guarded(c, [&] {
        if (sgs_order != CFDG_SGS_MULTICOLOUR && sgs_order != CFDG_SGS_SEQUENTIAL)
            throw CfdgError(CFDG_ERR_CONFIG, "unknown SGS order");
        if (sgs_order == CFDG_SGS_SEQUENTIAL) build_seq_groups(*c);
        const int gf = (sgs_order == CFDG_SGS_SEQUENTIAL) ? 2 : 0;
        for (int sweep = 0; sweep < 2; ++sweep) {
            const auto& off = c->h_gofs[gf + sweep];
            int32_t* out = sweep == 0 ? fwd_group : bwd_group;
            if (n_groups) n_groups[sweep] = static_cast<int32_t>(off.size()) - 1;
            if (!out) continue;
            // the rows the kernels are handed, read back from the device
            std::vector<i32> rows(c->grows[gf + sweep].size());
            c->d2h(rows.data(), c->grows[gf + sweep].get(), rows.size());
            c->sync();
            std::fill(out, out + c->nc, -1);
            for (std::size_t g = 0; g + 1 < off.size(); ++g)
                for (i32 k = off[g]; k < off[g + 1]; ++k) out[c->new2old[z(rows[z(k)])]] = static_cast<int32_t>(g);
        }
    });
}

int cfdg_spmv(cfdg_ctx* c, const double* x, double* y) {
    return guarded(c, [&] {
        if (!x || !y) throw CfdgError(CFDG_ERR_ARG, "null vector");
        require_jacobian(*c);
        const std::size_t n5 = 5ULL * z(c->nc);
        const auto d = c->vec_to_device(x);
        c->h2d(c->s.get(), d.data(), n5);
        spmv_stored<double>(*c, c->s.get(), c->t.get());
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
        cfdg_step_result st{};
        implicit_step(*c, *lp, &st);
        if (out) *out = st;
    });
}

int cfdg_explicit_step(cfdg_ctx* c, cfdg_step_result* out) {
    return guarded(c, [&] {
        if (c->sa_on()) throw CfdgError(CFDG_ERR_CONFIG, "explicit stepping with the turbulence module is not supported");
        if (c->bdf_order > 0) throw CfdgError(CFDG_ERR_STATE, "explicit stepping in dual time is not supported");
        eval_residual(*c);
        cfdg::launch_dt_explicit(c->margs(), c->fargs(), c->lam.get(), c->num.cfl, c->dt.get(), c->alpha.get(),
                                 c->stream);
        c->have_dt = false;  // alpha holds the explicit dt / V, not the implicit diagonal shift
        double l2[5];
        residual_norms(*c, l2);
        cfdg::launch_explicit_update(c->margs(), c->fargs(), c->alpha.get(), c->U.get(), c->stream);
        c->check_launch();
        c->sync();
        c->have_residual = false;
        if (out) {
            *out = cfdg_step_result{};
            for (int v = 0; v < 5; ++v) out->l2[v] = l2[v];
        }
    });
}

int cfdg_set_unsteady(cfdg_ctx* c, double dt, int32_t order) {
    return guarded(c, [&] {
        if (order < 0 || order > 2 || (order > 0 && !(dt > 0.0)))
            throw CfdgError(CFDG_ERR_ARG, "dual time stepping needs order 0, 1 or 2 and dt > 0");
        if (order > 0 && c->sa_on() && c->turb.update != CFDG_TURB_FROZEN)
            throw CfdgError(CFDG_ERR_CONFIG, "dual time stepping supports the turbulence module only with the frozen update");
        c->bdf_order = order;
        c->phys_dt = dt;
        c->bdf_cur = 1;
        c->phys_step = 0;
        const std::size_t n = 5 * z(c->nc);
        if (order > 0) {
            if (c->Un.size() != n) c->Un.alloc(n);
            if (c->Unm1.size() != n) c->Unm1.alloc(n);
            hip_check(hipMemcpyAsync(c->Un.get(), c->U.get(), n * sizeof(double), hipMemcpyDeviceToDevice, c->stream),
                      "copy");
            hip_check(hipMemcpyAsync(c->Unm1.get(), c->U.get(), n * sizeof(double), hipMemcpyDeviceToDevice, c->stream),
                      "copy");
            c->sync();
        }
        c->have_residual = false;
        c->have_dt = false;
    });
}

int cfdg_run_unsteady(cfdg_ctx* c, const cfdg_linparams* lp, int64_t steps, int64_t max_subiter, double sub_tol,
                      int64_t* subiters) {
    return guarded(c, [&] {
        if (!lp) throw CfdgError(CFDG_ERR_ARG, "null linear parameters");
        if (c->bdf_order == 0) throw CfdgError(CFDG_ERR_STATE, "dual time stepping is not enabled");
        if (steps < 0 || max_subiter < 1 || !(sub_tol >= 0.0))
            throw CfdgError(CFDG_ERR_ARG, "steps >= 0, max_subiter >= 1 and sub_tol >= 0 required");
        const std::size_t n = 5 * z(c->nc);
        for (int64_t k = 0; k < steps; ++k) {
            const int64_t step = ++c->phys_step;  // counts on across calls
            c->bdf_cur = (step == 1 || c->bdf_order == 1) ? 1 : 2;
            double norm0[5] = {0, 0, 0, 0, 0};
            bool have0 = false;
            int64_t last = 0;
            for (int64_t it = 1; it <= max_subiter; ++it) {
                last = it;
                cfdg_step_result st{};
                implicit_step(*c, *lp, &st);
                const bool diagnose = it == 1 || it % 5 == 0 || it == max_subiter;
                if (!diagnose) continue;
                if (!std::isfinite(st.l2[0])) throw CfdgError(CFDG_ERR_STATE, "non-finite residual in dual time stepping");
                double rel = 1.0;
                if (!have0) {
                    for (int v = 0; v < 5; ++v) norm0[v] = st.l2[v];
                    have0 = true;
                } else {
                    rel = 0.0;
                    for (int v = 0; v < 5; ++v) rel = std::max(rel, st.l2[v] / std::max(norm0[v], 1.0e-300));
                }
                if (rel <= sub_tol) break;
            }
            if (subiters) subiters[k] = last;
            // U^{n-1} = U^n, U^n = U^{n+1}
            if (c->bdf_order == 2) c->Unm1.swap(c->Un);
            hip_check(hipMemcpyAsync(c->Un.get(), c->U.get(), n * sizeof(double), hipMemcpyDeviceToDevice, c->stream),
                      "copy");
            c->sync();
        }
    });
}

int cfdg_boundary_integrals(cfdg_ctx* c, double* mass, double* energy, double* force, int32_t n_patches) {
    return guarded(c, [&] {
        if (!c->have_residual) throw CfdgError(CFDG_ERR_STATE, "boundary_integrals needs an evaluated residual");
        if (n_patches != c->npatch) throw CfdgError(CFDG_ERR_ARG, "n_patches must equal the mesh's patch count");
        cfdg::launch_boundary_integrals(c->margs(), c->fargs(), c->npatch, c->bint.get(), c->stream);
        c->check_launch();
        std::vector<double> h(5 * static_cast<std::size_t>(c->npatch));
        c->d2h(h.data(), c->bint.get(), h.size());
        c->sync();
        for (int p = 0; p < c->npatch; ++p) {
            const double* o = h.data() + 5 * p;
            if (mass) mass[p] = o[0];
            if (energy) energy[p] = o[4];
            if (force)
                for (int k = 0; k < 3; ++k) force[3 * p + k] = o[1 + k];
        }
    });
}

int cfdg_set_turbulence(cfdg_ctx* c, const cfdg_mesh* mesh, const cfdg_turbulence* t) {
    return guarded(c, [&] {
        if (!t) throw CfdgError(CFDG_ERR_ARG, "null turbulence parameters");
        // in dual time stepping the module block is copied unchanged, as the
        // reference's implicit Euler does (implicit_euler.hpp:64-72)
        if (t->model != CFDG_TURB_NONE && c->bdf_order > 0 && t->update != CFDG_TURB_FROZEN)
            throw CfdgError(CFDG_ERR_CONFIG, "dual time stepping supports the turbulence module only with the frozen update");
        c->have_residual = false;
        c->have_dt = false;
        if (t->model == CFDG_TURB_NONE) {
            c->turb = *t;
            return;
        }
        if (t->model != CFDG_TURB_SA) throw CfdgError(CFDG_ERR_CONFIG, "unsupported turbulence model");
        c->mut_ext = false;  // the module forms its own eddy viscosity
        c->sa_evaluated = false;
        if (t->update != CFDG_TURB_FROZEN && t->update != CFDG_TURB_POINT_IMPLICIT)
            throw CfdgError(CFDG_ERR_CONFIG, "turbulence update must be FROZEN or POINT_IMPLICIT");
        // config.cpp:175-186
        if (c->num.flow_model != CFDG_VISCOUS_FLOW)
            throw CfdgError(CFDG_ERR_CONFIG, "turbulence requires the viscous flow model");
        if (!(t->nu_inf_ratio > 0.0) || !(t->distance_tolerance > 0.0) || t->max_distance_sweeps < 0 ||
            !(t->rho_inf > 0.0) || !(t->p_inf > 0.0))
            throw CfdgError(CFDG_ERR_ARG, "nu_inf_ratio, distance_tolerance, rho_inf, p_inf > 0 and sweeps >= 0 required");
        if (!mesh) throw CfdgError(CFDG_ERR_ARG, "null mesh");
        const cfdg::Mesh& m = mesh->m;
        if (m.n_cells != c->nc || m.n_faces != c->nf || m.n_inner_faces != c->ninner)
            throw CfdgError(CFDG_ERR_ARG, "mesh differs from the mesh of cfdg_create");
        // zone treatment per boundary face (spalart_allmaras.hpp:167-181, 447-464)
        std::vector<char> wall(z(c->npatch), 0);
        std::vector<int> zone(z(c->npatch), cfdg::kZoneSymmetry);
        for (int p = 0; p < c->npatch; ++p) {
            switch (c->bcs.bc[p].type) {
                case cfdg::kNoSlip:
                case cfdg::kNoSlipHF: zone[z(p)] = cfdg::kZoneWall; break;
                case cfdg::kSupIn:
                case cfdg::kSubIn:
                case cfdg::kFar: zone[z(p)] = cfdg::kZoneInflow; break;
                case cfdg::kSupOut:
                case cfdg::kSubOut: zone[z(p)] = cfdg::kZoneOutflow; break;
                default: break;
            }
            wall[z(p)] = zone[z(p)] == cfdg::kZoneWall;
        }
        std::vector<int> kind(z(c->nbf));
        for (int b = 0; b < c->nbf; ++b) {
            const int p = m.face_patch[z(c->ninner + b)];
            kind[z(b)] = (p >= 0 && p < c->npatch) ? zone[z(p)] : cfdg::kZoneSymmetry;
        }
        // wall distance over the vertex stencil (solve_wall_distance, spalart_allmaras.hpp:183-189)
        const bool own_stencil = !(mesh->conn && mesh->conn->has_node_stencil);
        const cfdg::Csr st = own_stencil ? cfdg::cell_cells_by_node(m) : cfdg::Csr{};
        c->h_sa_dist = cfdg::wall_distance(m, own_stencil ? st : mesh->conn->cell_cells_node, wall,
                                           t->max_distance_sweeps, t->distance_tolerance);
        std::vector<double> dist_dev(z(c->nc));
        for (int r = 0; r < c->nc; ++r) dist_dev[z(r)] = c->h_sa_dist[z(c->new2old[z(r)])];
        // freestream working values (spalart_allmaras.hpp:191-193)
        const double T_inf = t->p_inf / (t->rho_inf * c->num.gas_constant);
        c->sa_nu_inf = t->nu_inf_ratio * cfdg::sutherland(T_inf) / t->rho_inf;
        c->sa_rho_nu_inf = t->rho_inf * c->sa_nu_inf;
        c->sa_kind.upload(kind, c->stream);
        c->sa_dist.upload(dist_dev, c->stream);
        c->sa_rnu.upload(std::vector<double>(z(c->ntot), c->sa_rho_nu_inf), c->stream);
        c->sa_nu.alloc(z(c->ntot));
        c->sa_rho.alloc(z(c->ntot));
        c->sa_mut.alloc(z(c->ntot));
        c->sa_mut.zero(c->stream);
        c->sa_g.alloc(3 * z(c->nc));
        c->sa_res.alloc(z(c->nc));
        c->sa_diag.alloc(z(c->nc));
        c->sa_mdot.alloc(z(c->nf));
        c->sa_mdot_tile.alloc(z(c->tld));
        c->sync();
        c->turb = *t;
    });
}

int cfdg_set_turbulence_state(cfdg_ctx* c, const double* rho_nu) {
    return guarded(c, [&] {
        if (!c->sa_on()) throw CfdgError(CFDG_ERR_STATE, "no turbulence model set");
        if (!rho_nu) throw CfdgError(CFDG_ERR_ARG, "null state");
        std::vector<double> h(z(c->nc));
        for (int r = 0; r < c->nc; ++r) h[z(r)] = rho_nu[c->new2old[z(r)]];
        c->h2d(c->sa_rnu.get(), h.data(), h.size());
        c->sync();
        c->have_residual = false;
        c->have_dt = false;
    });
}

int cfdg_set_eddy_viscosity(cfdg_ctx* c, const double* mut, int32_t n_total) {
    return guarded(c, [&] {
        c->have_residual = false;
        c->have_dt = false;
        if (!mut) {
            c->mut_ext = false;
            return;
        }
        if (c->sa_on()) throw CfdgError(CFDG_ERR_STATE, "the turbulence module forms its own eddy viscosity");
        if (c->num.flow_model != CFDG_VISCOUS_FLOW)
            throw CfdgError(CFDG_ERR_CONFIG, "eddy viscosity requires the viscous flow model");
        if (n_total != c->ntot) throw CfdgError(CFDG_ERR_ARG, "eddy viscosity needs n_total values");
        // cells to device order; ghosts keep their index n_cells + boundary face
        std::vector<double> h(z(c->ntot));
        for (int r = 0; r < c->nc; ++r) h[z(r)] = mut[c->new2old[z(r)]];
        for (int g = c->nc; g < c->ntot; ++g) h[z(g)] = mut[g];
        if (c->sa_mut.size() != z(c->ntot)) c->sa_mut.alloc(z(c->ntot));
        if (c->sa_mdot_tile.size() != z(c->tld)) c->sa_mdot_tile.alloc(z(c->tld));
        c->h2d(c->sa_mut.get(), h.data(), h.size());
        c->sync();
        c->mut_ext = true;
    });
}

int cfdg_get_turbulence(cfdg_ctx* c, double* rho_nu, double* nu, double* mut, double* dist, double* res,
                        double* grad, double* mdot, double* diag) {
    return guarded(c, [&] {
        if (!c->sa_on()) throw CfdgError(CFDG_ERR_STATE, "no turbulence model set");
        const auto cells = [&](const DBuf<double>& b, double* out, long long off) {
            if (!out) return;
            std::vector<double> h(z(c->nc));
            c->d2h(h.data(), b.get() + off, h.size());
            c->sync();
            for (int r = 0; r < c->nc; ++r) out[c->new2old[z(r)]] = h[z(r)];
        };
        cells(c->sa_rnu, rho_nu, 0);
        cells(c->sa_nu, nu, 0);
        cells(c->sa_mut, mut, 0);
        cells(c->sa_res, res, 0);
        cells(c->sa_diag, diag, 0);
        if (dist) std::copy(c->h_sa_dist.begin(), c->h_sa_dist.end(), dist);
        if (grad)
            for (int d = 0; d < 3; ++d) cells(c->sa_g, grad + static_cast<long long>(d) * c->nc, static_cast<long long>(d) * c->nc);
        if (mdot) {
            cfdg::launch_gather_mdot(c->nf, c->face_entry.get(), c->sa_mdot_tile.get(), c->sa_mdot.get(), c->stream);
            c->check_launch();
            c->d2h(mdot, c->sa_mdot.get(), z(c->nf));
            c->sync();
        }
    });
}

int cfdg_run_steady(cfdg_ctx* c, const cfdg_linparams* lp, const cfdg_steady_params* sp, cfdg_steady_result* out) {
    return guarded(c, [&] {
        if (!lp || !sp || !out) throw CfdgError(CFDG_ERR_ARG, "null argument");
        if (sp->max_steps < 1 || sp->check_interval < 1 || !(sp->rel_tol >= 0.0))
            throw CfdgError(CFDG_ERR_ARG, "max_steps >= 1, check_interval >= 1 and rel_tol >= 0 required");
        *out = cfdg_steady_result{};
        out->status = CFDG_MAX_ITERATIONS;
        const auto t0 = std::chrono::steady_clock::now();
        bool have0 = false;
        for (int64_t it = 1; it <= sp->max_steps; ++it) {
            cfdg_step_result st{};
            implicit_step(*c, *lp, &st);
            out->steps = it;
            out->linear_iterations += st.linear.iterations;
            const bool check = it == 1 || it % sp->check_interval == 0 || it == sp->max_steps;
            if (!check) continue;
            if (!std::isfinite(st.l2[0])) {
                out->status = CFDG_DIVERGED;
                break;
            }
            for (int v = 0; v < 5; ++v) out->l2_last[v] = st.l2[v];
            if (!have0) {
                for (int v = 0; v < 5; ++v) out->l2_first[v] = st.l2[v];
                have0 = true;
                out->rel = 1.0;
                continue;
            }
            double rel = 0.0;
            for (int v = 0; v < 5; ++v) rel = std::max(rel, st.l2[v] / std::max(out->l2_first[v], 1.0e-300));
            out->rel = rel;
            if (rel <= sp->rel_tol) {
                out->status = CFDG_CONVERGED;
                break;
            }
        }
        out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
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

int cfdg_gradient_form(const cfdg_ctx* c, int32_t* form, int32_t* n_gradient_tiles, int32_t* n_flux_tiles) {
    if (!c) return CFDG_ERR_ARG;
    const bool lds = c->tile_halo_off.size() > 0;
    if (form) *form = !lds ? CFDG_GRAD_GATHER : (c->gslots > cfdg::kGradSlots ? CFDG_GRAD_LDS_WIDE : CFDG_GRAD_LDS);
    if (n_gradient_tiles) *n_gradient_tiles = lds ? c->ngtiles : 0;
    if (n_flux_tiles) *n_flux_tiles = c->ntiles;
    return CFDG_OK;
}

int cfdg_residual_norms_noalloc(const double* res, int64_t ld, int32_t n_cells, double* l2, void* stream) {
    if (!res || !l2 || n_cells < 0) return CFDG_ERR_ARG;
    if (n_cells > CFDG_NOALLOC_MAX_CELLS) return CFDG_ERR_CAPACITY;
    if (ld < n_cells || ld > static_cast<int64_t>(INT32_MAX)) return CFDG_ERR_ARG;
    const hipError_t e = cfdg::residual_norms_noalloc(n_cells, static_cast<int>(ld), res, l2,
                                                      static_cast<hipStream_t>(stream));
    return e == hipSuccess ? CFDG_OK : CFDG_ERR_HIP;
}

}  // extern "C"
