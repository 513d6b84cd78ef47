// Device context of the C ABI (include/cfdg.h): device buffers, stream,
// permutations and the state of the precision forms and the learned warm
// start.  Shared by the C-ABI translation units; not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "cfdg.h"
#include "kernels.hpp"
#include "setup.hpp"

namespace cfdg {

inline std::size_t z(long long v) { return static_cast<std::size_t>(v); }

inline void hip_check(hipError_t e, const char* what) {
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
    void swap(DBuf& o) noexcept {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
    }

  private:
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* p_ = nullptr;
    std::size_t n_ = 0;
};


}  // namespace cfdg

struct cfdg_ctx;

namespace cfdg {
// Learned warm start hooks of the implicit step (capi_gnn.cpp).
// Features of the current step (after the residual and dt): when recording or
// when a network is attached.
void gnn_step_features(cfdg_ctx& c);
// Writes the attached network's prediction of the update into x0 (device
// order, AoS); false when no network is attached.
bool gnn_warm_guess(cfdg_ctx& c, double* x0);
}  // namespace cfdg

struct cfdg_ctx {
    template <typename T>
    using DBuf = cfdg::DBuf<T>;
    static void hip_check(hipError_t e, const char* what) { cfdg::hip_check(e, what); }
    static std::size_t z(long long v) { return cfdg::z(v); }

    std::string err;
    hipStream_t stream = nullptr;
    int nc = 0, nf = 0, ninner = 0, nbf = 0, ntot = 0, npatch = 0, ntiles = 0;
    cfdg_numerics num{};
    cfdg::NumArgs nargs{};
    cfdg::BcTable bcs{};
    std::vector<std::int32_t> new2old, old2new;

    // mesh (device)
    DBuf<double> fp, tf;          // face planes: mesh order (fld), tile face-list order (tld)
    DBuf<int2> fcells;            // (owner, neighbour or ghost) per face
    long long fld = 0, tld = 0;
    DBuf<int> cf_off, cf, cf_slot, ccf_off, ccf, ccn_off, ccn, bface_bc, tile_off, tile_face_off;
    DBuf<int> tile_halo_off, tile_halo, gtile_off;  // gradient tiles (the flux tiles or sub-ranges of them)
    int ngtiles = 0, gblock = 0, gslots = 0;
    DBuf<std::uint16_t> ccn_slot, ccf_slot;
    DBuf<int> ccn_base, ccf_base;
    DBuf<double> face_vec;        // limiter face vectors per gradient tile (transposed)
    DBuf<long long> face_vec_base;
    // fields (device order)
    DBuf<double> U, cs, res, lam, dt, alpha, bflux, bint;
    long long ld = 0;  // cell plane stride
    bool have_residual = false;
    bool have_dt = false;
    bool have_jacobian = false;  // block values set (assembled, or given) in the current storage
    // BSR: mesh-order pattern for the ABI, device-order pattern for the kernels
    std::vector<std::int32_t> h_row_ptr, h_cols, h_diag;
    std::vector<long long> blk_dev;  // mesh block k -> device block
    DBuf<int> row_ptr, cols, diag, cf_blk;
    // per block: 1 when its column lies in an earlier group of the forward SGS
    // sweep than its row ([0] multicolour groups, [1] dependency levels)
    DBuf<std::uint8_t> blk_lower[2];
    // per device row: 1 when it lies in the forward sweep's first group (same slots)
    DBuf<std::uint8_t> row_first[2];
    // block values and diagonal LU as raw bytes in the storage of the precision form
    DBuf<std::uint8_t> values, lu;
    int storage = cfdg::kStoreF64;
    int lu_f32 = 0;
    cfdg_precision prec{CFDG_PREC_FP64, CFDG_PREC_FP64, 8, 0, 0.1};
    DBuf<std::uint8_t> piv;
    DBuf<int> singular;
    // SGS row groups (device rows): 0 multicolour fwd, 1 multicolour bwd, 2 sequential fwd, 3 sequential bwd
    std::vector<std::int32_t> h_gofs[4];
    DBuf<int> grows[4];
    cfdg::Csr ccf_mesh;
    bool have_seq_groups = false;
    // Krylov work vectors (AoS, 5 nc, device order); FP32 set for the FP32 form
    DBuf<double> b, x, r, r0, p, v, s, t, ph, sh;
    struct F32 {
        DBuf<float> b, x, r, r0, p, v, s, t, ph, sh;
    } f32;
    DBuf<double> partial, dots;
    double* h_dots = nullptr;  // pinned
    DBuf<double> ks;           // device-driven BiCGSTAB state block (kernels_linalg.hpp)
    double* h_ks = nullptr;    // pinned copy
    int last_iters[2] = {4, 4};  // iterations of the last FP64 / FP32 solve: the first batch size
    // learned warm start (gnn.hpp)
    std::uint64_t uid = 0;              // identity of this context's graph in recorded samples
    bool have_ref = false;              // freestream of the feature normalisation
    double ref_rho = 0.0, ref_vel[3] = {0.0, 0.0, 0.0}, ref_p = 0.0;
    bool gnn_record = false;            // keep every step's features for cfdg_gnn_record
    struct cfdg_gnn* warm = nullptr;    // network of the initial guess (null: zero guess)
    bool have_graph = false;
    DBuf<int> gadj, grev;
    DBuf<float> geg;
    double h_mean = 0.0;
    DBuf<cfdg::bf16> gx;                // features of the current step
    DBuf<double> gnorm;                 // [0] squared norm of -R / lambda, [1] the previous step's, then partials
    bool have_prev_du = false;          // x holds the previous implicit step's update
    DBuf<double> du_prev;               // the previous step's update, copied when the features are formed
    bool features_had_prev = false;     // du_prev was valid for the current features
    bool have_features = false;
    long long warm_used = 0, warm_rejected = 0;
    // dual time stepping (off while bdf_order == 0): physical time step, the
    // BDF order of the current physical step and the time levels U^n, U^{n-1}
    int bdf_order = 0, bdf_cur = 1;
    int64_t phys_step = 0;
    double phys_dt = 0.0;
    DBuf<double> Un, Unm1;
    // Spalart-Allmaras module (off unless turb.model == CFDG_TURB_SA)
    cfdg_turbulence turb{};
    double sa_nu_inf = 0.0, sa_rho_nu_inf = 0.0;
    DBuf<int> sa_kind;                                  // zone of each boundary face
    DBuf<double> sa_dist, sa_rnu, sa_nu, sa_rho, sa_mut, sa_g, sa_res, sa_diag, sa_mdot;
    DBuf<double> sa_mdot_tile;                          // mass flux per tile face-list entry
    DBuf<int> face_entry;                               // entry of each face in its owner tile's face list
    std::vector<double> h_sa_dist;                      // wall distance, mesh order
    bool sa_on() const { return turb.model == CFDG_TURB_SA; }
    // prescribed eddy viscosity (cfdg_set_eddy_viscosity): sa_mut holds it and
    // sa_mdot_tile takes the flux kernel's mass flux; the module stays off
    bool mut_ext = false;
    bool sa_evaluated = false;  // sa_nu / sa_mut hold a residual evaluation's values
    bool has_mut() const { return sa_on() || mut_ext; }
    cfdg::SaArgs sargs() const {
        cfdg::SaArgs a{};
        a.nu_inf = sa_nu_inf;
        a.kind = sa_kind.get();
        a.dist = sa_dist.get();
        a.rnu = sa_rnu.get();
        a.nu = sa_nu.get();
        a.rho = sa_rho.get();
        a.mut = sa_mut.get();
        a.g = sa_g.get();
        a.res = sa_res.get();
        a.diag = sa_diag.get();
        a.mdot = sa_mdot.get();
        a.mdot_tile = sa_mdot_tile.get();
        return a;
    }
    // timing
    hipEvent_t ev[8] = {};
    double stage_ms[6] = {0, 0, 0, 0, 0, 0};

    cfdg::MeshArgs margs() const {
        cfdg::MeshArgs a{};
        a.nc = nc;
        a.nf = nf;
        a.ninner = ninner;
        a.ntiles = ntiles;
        a.fp = fp.get();
        a.fld = fld;
        a.fcells = fcells.get();
        a.tf = tf.get();
        a.tld = tld;
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
        a.tile_halo_off = tile_halo_off.get();
        a.gtile_off = gtile_off.get();
        a.ngtiles = ngtiles;
        a.gblock = gblock;
        a.gslots = gslots;
        a.tile_halo = tile_halo.get();
        a.ccn_slot = ccn_slot.get();
        a.ccf_slot = ccf_slot.get();
        a.ccn_base = ccn_base.get();
        a.ccf_base = ccf_base.get();
        a.face_vec = face_vec.get();
        a.face_vec_base = face_vec_base.get();
        return a;
    }
    cfdg::FieldArgs fargs() const {
        cfdg::FieldArgs f{};
        f.U = U.get();
        f.cs = cs.get();
        f.ld = ld;
        f.res = res.get();
        f.lam = lam.get();
        f.bflux = bflux.get();
        f.mut = has_mut() ? sa_mut.get() : nullptr;
        f.mdot = sa_on() ? sa_mdot.get() : nullptr;
        f.mdot_tile = has_mut() ? sa_mdot_tile.get() : nullptr;
        return f;
    }
    cfdg::BsrArgs bargs() const {
        cfdg::BsrArgs a{};
        a.n = nc;
        a.row_ptr = row_ptr.get();
        a.cols = cols.get();
        a.diag = diag.get();
        a.cf_blk = cf_blk.get();
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
    // all cell planes (kCellPlanes x ld), device order
    std::vector<double> download_planes() {
        std::vector<double> h(z(ld) * cfdg::kCellPlanes);
        d2h(h.data(), cs.get(), h.size());
        sync();
        return h;
    }

    ~cfdg_ctx() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (h_dots) (void)hipHostFree(h_dots);
        if (h_ks) (void)hipHostFree(h_ks);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

