// Mesh container reader / writer (container.hpp), libhdf5 loaded at run time.
#include "container.hpp"

#include <dlfcn.h>
#include <hdf5.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace cfdg::front {

namespace {

[[noreturn]] void fail(const std::string& what) { throw ContainerError(what); }

// H5F_ACC_RDONLY / H5F_ACC_TRUNC as values: the header's macros also call
// H5check() and H5open(), which would make libhdf5 a link dependency.
constexpr unsigned kAccRdonly = 0x0000u, kAccTrunc = 0x0002u;

// The subset of the HDF5 C API used here, resolved from the shared library.
struct Hdf5 {
    herr_t (*open)();
    hid_t (*Fopen)(const char*, unsigned, hid_t);
    hid_t (*Fcreate)(const char*, unsigned, hid_t, hid_t);
    herr_t (*Fclose)(hid_t);
    hid_t (*Gopen2)(hid_t, const char*, hid_t);
    hid_t (*Gcreate2)(hid_t, const char*, hid_t, hid_t, hid_t);
    herr_t (*Gclose)(hid_t);
    hid_t (*Dopen2)(hid_t, const char*, hid_t);
    hid_t (*Dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
    herr_t (*Dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
    herr_t (*Dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
    hid_t (*Dget_space)(hid_t);
    herr_t (*Dclose)(hid_t);
    int (*Sget_simple_extent_ndims)(hid_t);
    int (*Sget_simple_extent_dims)(hid_t, hsize_t*, hsize_t*);
    hid_t (*Screate_simple)(int, const hsize_t*, const hsize_t*);
    hid_t (*Screate)(H5S_class_t);
    herr_t (*Sclose)(hid_t);
    hid_t (*Aopen)(hid_t, const char*, hid_t);
    hid_t (*Acreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t);
    herr_t (*Aread)(hid_t, hid_t, void*);
    herr_t (*Awrite)(hid_t, hid_t, const void*);
    hid_t (*Aget_space)(hid_t);
    herr_t (*Aclose)(hid_t);
    hid_t (*Tcopy)(hid_t);
    herr_t (*Tset_size)(hid_t, size_t);
    herr_t (*Tclose)(hid_t);
    herr_t (*Eset_auto2)(hid_t, H5E_auto2_t, void*);
    htri_t (*Lexists)(hid_t, const char*, hid_t);
    // native types (globals, valid after H5open)
    hid_t f64, u8, i32, i64, u64, native_int, c_s1;
};

Hdf5 g_h5{};
std::string g_h5_error;
bool g_h5_loaded = false;
std::mutex g_h5_mutex;

template <class F>
void bind(void* lib, const char* name, F& fn) {
    void* p = dlsym(lib, name);
    if (!p) throw ContainerError(std::string("libhdf5: missing symbol ") + name);
    fn = reinterpret_cast<F>(p);
}
hid_t global_id(void* lib, const char* name) {
    void* p = dlsym(lib, name);
    if (!p) throw ContainerError(std::string("libhdf5: missing symbol ") + name);
    return *static_cast<const hid_t*>(p);
}

const Hdf5& h5() {
    std::lock_guard<std::mutex> lk(g_h5_mutex);
    if (g_h5_loaded) return g_h5;
    if (!g_h5_error.empty()) fail(g_h5_error);
    std::vector<std::string> names;
    if (const char* e = std::getenv("CFDG_HDF5_LIB")) names.emplace_back(e);
    names.insert(names.end(), {"/opt/conda/lib/libhdf5.so.103", "libhdf5.so.103", "libhdf5.so"});
    void* lib = nullptr;
    for (const auto& n : names)
        if ((lib = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL)) != nullptr) break;
    if (!lib) {
        g_h5_error = "cannot load libhdf5 (set CFDG_HDF5_LIB to the library's path)";
        fail(g_h5_error);
    }
    try {
        Hdf5 h{};
        bind(lib, "H5open", h.open);
        bind(lib, "H5Fopen", h.Fopen);
        bind(lib, "H5Fcreate", h.Fcreate);
        bind(lib, "H5Fclose", h.Fclose);
        bind(lib, "H5Gopen2", h.Gopen2);
        bind(lib, "H5Gcreate2", h.Gcreate2);
        bind(lib, "H5Gclose", h.Gclose);
        bind(lib, "H5Dopen2", h.Dopen2);
        bind(lib, "H5Dcreate2", h.Dcreate2);
        bind(lib, "H5Dread", h.Dread);
        bind(lib, "H5Dwrite", h.Dwrite);
        bind(lib, "H5Dget_space", h.Dget_space);
        bind(lib, "H5Dclose", h.Dclose);
        bind(lib, "H5Sget_simple_extent_ndims", h.Sget_simple_extent_ndims);
        bind(lib, "H5Sget_simple_extent_dims", h.Sget_simple_extent_dims);
        bind(lib, "H5Screate_simple", h.Screate_simple);
        bind(lib, "H5Screate", h.Screate);
        bind(lib, "H5Sclose", h.Sclose);
        bind(lib, "H5Aopen", h.Aopen);
        bind(lib, "H5Acreate2", h.Acreate2);
        bind(lib, "H5Aread", h.Aread);
        bind(lib, "H5Awrite", h.Awrite);
        bind(lib, "H5Aget_space", h.Aget_space);
        bind(lib, "H5Aclose", h.Aclose);
        bind(lib, "H5Tcopy", h.Tcopy);
        bind(lib, "H5Tset_size", h.Tset_size);
        bind(lib, "H5Tclose", h.Tclose);
        bind(lib, "H5Eset_auto2", h.Eset_auto2);
        bind(lib, "H5Lexists", h.Lexists);
        if (h.open() < 0) fail("libhdf5: H5open failed");
        h.f64 = global_id(lib, "H5T_NATIVE_DOUBLE_g");
        h.u8 = global_id(lib, "H5T_NATIVE_UINT8_g");
        h.i32 = global_id(lib, "H5T_NATIVE_INT32_g");
        h.i64 = global_id(lib, "H5T_NATIVE_INT64_g");
        h.u64 = global_id(lib, "H5T_NATIVE_UINT64_g");
        h.native_int = global_id(lib, "H5T_NATIVE_INT_g");
        h.c_s1 = global_id(lib, "H5T_C_S1_g");
        h.Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // errors come back as return codes only
        g_h5 = h;
        g_h5_loaded = true;
    } catch (const ContainerError& e) {
        g_h5_error = e.what();
        throw;
    }
    return g_h5;
}

// Owning handle: closes with the function matching its kind.
class Id {
public:
    Id() = default;
    Id(hid_t id, herr_t (*close)(hid_t)) : id_(id), close_(close) {}
    Id(const Id&) = delete;
    Id& operator=(const Id&) = delete;
    Id(Id&& o) noexcept : id_(o.id_), close_(o.close_) { o.id_ = -1; }
    Id& operator=(Id&& o) noexcept {
        reset();
        id_ = o.id_;
        close_ = o.close_;
        o.id_ = -1;
        return *this;
    }
    ~Id() { reset(); }
    operator hid_t() const { return id_; }
    bool ok() const { return id_ >= 0; }

private:
    void reset() {
        if (id_ >= 0 && close_) close_(id_);
        id_ = -1;
    }
    hid_t id_ = -1;
    herr_t (*close_)(hid_t) = nullptr;
};

Id group_open(hid_t loc, const char* name) {
    const Hdf5& h = h5();
    Id g(h.Gopen2(loc, name, H5P_DEFAULT), h.Gclose);
    if (!g.ok()) fail(std::string("container: missing group /") + name);
    return g;
}
Id group_create(hid_t loc, const char* name) {
    const Hdf5& h = h5();
    Id g(h.Gcreate2(loc, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), h.Gclose);
    if (!g.ok()) fail(std::string("container: cannot create group /") + name);
    return g;
}

// dims of a dataset (rank 1 or 2)
std::vector<hsize_t> dataset_dims(hid_t d, const std::string& what) {
    const Hdf5& h = h5();
    Id sp(h.Dget_space(d), h.Sclose);
    const int nd = h.Sget_simple_extent_ndims(sp);
    if (nd < 0 || nd > 2) fail("container: " + what + " has an unexpected rank");
    std::vector<hsize_t> dims(static_cast<std::size_t>(nd));
    if (nd > 0 && h.Sget_simple_extent_dims(sp, dims.data(), nullptr) < 0) fail("container: cannot query " + what);
    return dims;
}

// the whole 1-D dataset grp/name as `type`, which must hold `expect` entries
// (expect < 0: any length)
template <class T>
std::vector<T> read_1d(hid_t grp, const char* gname, const char* name, hid_t type, long long expect) {
    const Hdf5& h = h5();
    const std::string what = std::string("/") + gname + "/" + name;
    Id d(h.Dopen2(grp, name, H5P_DEFAULT), h.Dclose);
    if (!d.ok()) fail("container: missing dataset " + what);
    const auto dims = dataset_dims(d, what);
    if (dims.size() != 1) fail("container: " + what + " is not one-dimensional");
    if (expect >= 0 && dims[0] != static_cast<hsize_t>(expect))
        fail("container: " + what + " holds " + std::to_string(dims[0]) + " entries, expected " +
             std::to_string(expect));
    std::vector<T> v(static_cast<std::size_t>(dims[0]));
    if (!v.empty() && h.Dread(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) < 0)
        fail("container: cannot read " + what);
    return v;
}

// [n, 3] doubles split into three arrays
void read_vec3(hid_t grp, const char* gname, const char* name, long long n, std::vector<double>& x,
               std::vector<double>& y, std::vector<double>& z) {
    const Hdf5& h = h5();
    const std::string what = std::string("/") + gname + "/" + name;
    Id d(h.Dopen2(grp, name, H5P_DEFAULT), h.Dclose);
    if (!d.ok()) fail("container: missing dataset " + what);
    const auto dims = dataset_dims(d, what);
    if (dims.size() != 2 || dims[1] != 3 || dims[0] != static_cast<hsize_t>(n))
        fail("container: " + what + " must be [" + std::to_string(n) + ", 3]");
    std::vector<double> flat(static_cast<std::size_t>(n) * 3);
    if (n > 0 && h.Dread(d, h.f64, H5S_ALL, H5S_ALL, H5P_DEFAULT, flat.data()) < 0)
        fail("container: cannot read " + what);
    x.resize(static_cast<std::size_t>(n));
    y.resize(static_cast<std::size_t>(n));
    z.resize(static_cast<std::size_t>(n));
    for (std::size_t i = 0; i < static_cast<std::size_t>(n); ++i) {
        x[i] = flat[3 * i];
        y[i] = flat[3 * i + 1];
        z[i] = flat[3 * i + 2];
    }
}

template <class T>
void read_attr(hid_t file, const char* name, hid_t type, T* out, std::size_t n) {
    const Hdf5& h = h5();
    Id a(h.Aopen(file, name, H5P_DEFAULT), h.Aclose);
    if (!a.ok()) fail(std::string("container: missing attribute ") + name);
    Id sp(h.Aget_space(a), h.Sclose);
    const int nd = h.Sget_simple_extent_ndims(sp);
    hsize_t dims[1] = {1};
    if (nd == 1) h.Sget_simple_extent_dims(sp, dims, nullptr);
    if ((nd == 0 && n != 1) || (nd == 1 && dims[0] != n) || nd > 1)
        fail(std::string("container: attribute ") + name + " has an unexpected shape");
    if (h.Aread(a, type, out) < 0) fail(std::string("container: cannot read attribute ") + name);
}

std::vector<std::string> read_strings(hid_t grp, const char* name, std::size_t n) {
    const Hdf5& h = h5();
    constexpr std::size_t kLen = 64;
    Id d(h.Dopen2(grp, name, H5P_DEFAULT), h.Dclose);
    if (!d.ok()) fail(std::string("container: missing dataset /patches/") + name);
    const auto dims = dataset_dims(d, name);
    if (dims.size() != 1 || dims[0] != n) fail(std::string("container: /patches/") + name + " has the wrong length");
    Id st(h.Tcopy(h.c_s1), h.Tclose);
    h.Tset_size(st, kLen);
    std::vector<char> buf(n * kLen, '\0');
    if (n > 0 && h.Dread(d, st, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0)
        fail(std::string("container: cannot read /patches/") + name);
    std::vector<std::string> out(n);
    for (std::size_t p = 0; p < n; ++p) out[p] = std::string(buf.data() + p * kLen, strnlen(buf.data() + p * kLen, kLen));
    return out;
}

// writers -------------------------------------------------------------------

template <class T>
void write_1d(hid_t grp, const char* name, hid_t type, const T* data, std::size_t n) {
    const Hdf5& h = h5();
    const hsize_t dims[1] = {static_cast<hsize_t>(n)};
    Id sp(h.Screate_simple(1, dims, nullptr), h.Sclose);
    Id d(h.Dcreate2(grp, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), h.Dclose);
    if (!d.ok()) fail(std::string("container: cannot create dataset ") + name);
    if (n > 0 && h.Dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
        fail(std::string("container: cannot write dataset ") + name);
}

void write_vec3(hid_t grp, const char* name, const std::vector<double>& x, const std::vector<double>& y,
                const std::vector<double>& z) {
    const Hdf5& h = h5();
    const std::size_t n = x.size();
    std::vector<double> flat(n * 3);
    for (std::size_t i = 0; i < n; ++i) {
        flat[3 * i] = x[i];
        flat[3 * i + 1] = y[i];
        flat[3 * i + 2] = z[i];
    }
    const hsize_t dims[2] = {static_cast<hsize_t>(n), 3};
    Id sp(h.Screate_simple(2, dims, nullptr), h.Sclose);
    Id d(h.Dcreate2(grp, name, h.f64, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), h.Dclose);
    if (!d.ok()) fail(std::string("container: cannot create dataset ") + name);
    if (n > 0 && h.Dwrite(d, h.f64, H5S_ALL, H5S_ALL, H5P_DEFAULT, flat.data()) < 0)
        fail(std::string("container: cannot write dataset ") + name);
}

// global ids: int32 when every value fits, else int64
void write_gid(hid_t grp, const char* name, const std::vector<std::int64_t>& g) {
    const Hdf5& h = h5();
    const bool fits = std::all_of(g.begin(), g.end(), [](std::int64_t v) {
        return v >= INT32_MIN && v <= INT32_MAX;
    });
    if (fits) {
        std::vector<std::int32_t> g32(g.begin(), g.end());
        write_1d(grp, name, h.i32, g32.data(), g32.size());
    } else {
        write_1d(grp, name, h.i64, g.data(), g.size());
    }
}

void write_strings(hid_t grp, const char* name, const std::vector<std::string>& s) {
    const Hdf5& h = h5();
    constexpr std::size_t kLen = 64;
    std::vector<char> buf(s.size() * kLen, '\0');
    for (std::size_t i = 0; i < s.size(); ++i) std::strncpy(buf.data() + i * kLen, s[i].c_str(), kLen - 1);
    Id st(h.Tcopy(h.c_s1), h.Tclose);
    h.Tset_size(st, kLen);
    const hsize_t dims[1] = {static_cast<hsize_t>(s.size())};
    Id sp(h.Screate_simple(1, dims, nullptr), h.Sclose);
    Id d(h.Dcreate2(grp, name, st, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), h.Dclose);
    if (!d.ok()) fail(std::string("container: cannot create dataset ") + name);
    if (!s.empty() && h.Dwrite(d, st, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0)
        fail(std::string("container: cannot write dataset ") + name);
}

}  // namespace

ContainerMesh read_container(const std::string& path) {
    const Hdf5& h = h5();
    Id file(h.Fopen(path.c_str(), kAccRdonly, H5P_DEFAULT), h.Fclose);
    if (!file.ok()) fail("Failed to open HDF5 mesh file for reading: " + path);
    ContainerMesh m;

    int nprocs = 0;
    read_attr(file, "nprocs", h.native_int, &nprocs, 1);
    if (nprocs != 1)
        fail("HDF5 mesh file was partitioned for " + std::to_string(nprocs) +
             " ranks, but current communicator has 1 ranks.");
    read_attr(file, "n_cells_global", h.u64, &m.n_cells_g, 1);
    read_attr(file, "n_faces_global", h.u64, &m.n_faces_g, 1);
    read_attr(file, "n_bfaces_global", h.u64, &m.n_bfaces_g, 1);
    read_attr(file, "n_nodes_global", h.u64, &m.n_nodes_g, 1);
    read_attr(file, "bbox_lo", h.f64, m.bbox_lo.data(), 3);
    read_attr(file, "bbox_hi", h.f64, m.bbox_hi.data(), 3);

    // partition tables: one partition, rank 0 -> partition rank2part[0]
    std::uint64_t c_cnt, cn_cnt, f_cnt, fn_cnt, n_cnt, nb_cnt;
    {
        Id g = group_open(file, "partition");
        const auto r2p = read_1d<int>(g, "partition", "rank2part", h.native_int, 1);
        if (r2p[0] != 0) fail("container: rank2part[0] must be 0 in a one-partition file");
        const auto span = [&](const char* name) {
            const auto o = read_1d<std::uint64_t>(g, "partition", name, h.u64, 2);
            if (o[0] != 0 || o[1] < o[0]) fail(std::string("container: bad /partition/") + name);
            return o[1];
        };
        c_cnt = span("cell_offsets");
        cn_cnt = span("cell_nodes_offsets");
        f_cnt = span("face_offsets");
        fn_cnt = span("face_nodes_offsets");
        n_cnt = span("node_offsets");
        nb_cnt = span("comm_nb_offsets");
    }
    if (c_cnt > static_cast<std::uint64_t>(INT32_MAX) || f_cnt > static_cast<std::uint64_t>(INT32_MAX) ||
        n_cnt > static_cast<std::uint64_t>(INT32_MAX) || cn_cnt > static_cast<std::uint64_t>(INT32_MAX) ||
        fn_cnt > static_cast<std::uint64_t>(INT32_MAX))
        fail("container: partition larger than 32-bit local indices");
    const auto nc = static_cast<long long>(c_cnt), nf = static_cast<long long>(f_cnt),
               nn = static_cast<long long>(n_cnt);

    {  // cells
        Id g = group_open(file, "cells");
        m.cell_type = read_1d<std::uint8_t>(g, "cells", "type", h.u8, nc);
        m.cell_gid = read_1d<std::int64_t>(g, "cells", "gid", h.i64, nc);
        m.cell_donor = read_1d<std::int32_t>(g, "cells", "donor", h.i32, nc);
        m.cell_volume = read_1d<double>(g, "cells", "volume", h.f64, nc);
        m.cell_nodes_off = read_1d<std::int32_t>(g, "cells", "nodes_offsets", h.i32, nc);
        m.cell_nodes_off.push_back(static_cast<std::int32_t>(cn_cnt));
        m.cell_nodes = read_1d<std::int32_t>(g, "cells", "nodes", h.i32, static_cast<long long>(cn_cnt));
        read_vec3(g, "cells", "centroid", nc, m.cell_cx, m.cell_cy, m.cell_cz);
        m.n_own = 0;
        while (m.n_own < nc && m.cell_donor[static_cast<std::size_t>(m.n_own)] == -1) ++m.n_own;
    }
    {  // faces
        Id g = group_open(file, "faces");
        m.face_owner = read_1d<std::int32_t>(g, "faces", "owner", h.i32, nf);
        m.face_neigh = read_1d<std::int32_t>(g, "faces", "neigh", h.i32, nf);
        m.face_patch = read_1d<std::int32_t>(g, "faces", "patch", h.i32, nf);
        m.face_type = read_1d<std::uint8_t>(g, "faces", "type", h.u8, nf);
        m.face_area = read_1d<double>(g, "faces", "area", h.f64, nf);
        m.face_nodes_off = read_1d<std::int32_t>(g, "faces", "nodes_offsets", h.i32, nf);
        m.face_nodes_off.push_back(static_cast<std::int32_t>(fn_cnt));
        m.face_nodes = read_1d<std::int32_t>(g, "faces", "nodes", h.i32, static_cast<long long>(fn_cnt));
        read_vec3(g, "faces", "normal", nf, m.face_nx, m.face_ny, m.face_nz);
        read_vec3(g, "faces", "centroid", nf, m.face_cx, m.face_cy, m.face_cz);
    }
    {  // nodes
        Id g = group_open(file, "nodes");
        m.node_gid = read_1d<std::int64_t>(g, "nodes", "gid", h.i64, nn);
        read_vec3(g, "nodes", "coords", nn, m.node_x, m.node_y, m.node_z);
        std::int32_t mx = 0;
        for (std::int32_t c = 0; c < m.n_own; ++c)
            for (std::int32_t k = m.cell_nodes_off[static_cast<std::size_t>(c)];
                 k < m.cell_nodes_off[static_cast<std::size_t>(c) + 1]; ++k)
                mx = std::max(mx, m.cell_nodes[static_cast<std::size_t>(k)] + 1);
        m.n_nodes_own = mx;
    }
    {  // comm: one partition has no neighbours
        Id g = group_open(file, "comm");
        m.nb_ranks = read_1d<std::int32_t>(g, "comm", "nb_ranks", h.i32, static_cast<long long>(nb_cnt));
        m.send_offsets = read_1d<std::int32_t>(g, "comm", "send_offsets", h.i32, static_cast<long long>(nb_cnt) + 1);
        m.recv_offsets = read_1d<std::int32_t>(g, "comm", "recv_offsets", h.i32, static_cast<long long>(nb_cnt) + 1);
        m.send_owned = read_1d<std::int32_t>(g, "comm", "send_owned_cells", h.i32, m.send_offsets.back());
        m.recv_ghost = read_1d<std::int32_t>(g, "comm", "recv_ghost_cells", h.i32, m.recv_offsets.back());
    }
    {  // patches
        Id g = group_open(file, "patches");
        Id d(h.Dopen2(g, "names", H5P_DEFAULT), h.Dclose);
        if (!d.ok()) fail("container: missing dataset /patches/names");
        const auto dims = dataset_dims(d, "/patches/names");
        if (dims.size() != 1) fail("container: /patches/names is not one-dimensional");
        const std::size_t np = dims[0];
        m.patch_names = read_strings(g, "names", np);
        m.patch_cgns_types = read_strings(g, "cgns_types", np);
        m.patch_face_offsets =
            read_1d<std::int32_t>(g, "patches", "patch_face_offsets", h.i32, static_cast<long long>(np) + 1);
        m.patch_faces = read_1d<std::int32_t>(g, "patches", "patch_faces", h.i32, m.patch_face_offsets.back());
    }
    // 3-zone face order: interior faces first (hdf5_reader.cpp:421-450)
    for (std::int32_t f = 0; f < static_cast<std::int32_t>(nf); ++f) {
        const std::int32_t nb = m.face_neigh[static_cast<std::size_t>(f)];
        if (nb < 0) break;
        ++m.n_inner_faces;
        if (nb < m.n_own) {
            if (m.n_inner_faces != m.n_internal_faces + 1)
                fail("container: corrupted face layout (zone 0 must be contiguous at the start)");
            ++m.n_internal_faces;
        }
    }
    return m;
}

void write_container(const std::string& path, const ContainerMesh& m) {
    const Hdf5& h = h5();
    Id file(h.Fcreate(path.c_str(), kAccTrunc, H5P_DEFAULT, H5P_DEFAULT), h.Fclose);
    if (!file.ok()) fail("Failed to create HDF5 mesh file: " + path);
    const std::size_t nc = static_cast<std::size_t>(m.n_cells()), nf = static_cast<std::size_t>(m.n_faces());
    if (m.cell_nodes_off.size() != nc + 1 || m.face_nodes_off.size() != nf + 1 ||
        m.patch_face_offsets.size() != m.patch_names.size() + 1 || m.send_offsets.size() != m.nb_ranks.size() + 1 ||
        m.recv_offsets.size() != m.nb_ranks.size() + 1)
        fail("container: inconsistent mesh arrays");
    {  // root attributes
        Id scalar(h.Screate(H5S_SCALAR), h.Sclose);
        const auto attr = [&](const char* name, hid_t type, const void* v, hid_t space) {
            Id a(h.Acreate2(file, name, type, space, H5P_DEFAULT, H5P_DEFAULT), h.Aclose);
            if (!a.ok() || h.Awrite(a, type, v) < 0) fail(std::string("container: cannot write attribute ") + name);
        };
        const int one = 1;
        attr("nprocs", h.native_int, &one, scalar);
        attr("n_cells_global", h.u64, &m.n_cells_g, scalar);
        attr("n_faces_global", h.u64, &m.n_faces_g, scalar);
        attr("n_bfaces_global", h.u64, &m.n_bfaces_g, scalar);
        attr("n_nodes_global", h.u64, &m.n_nodes_g, scalar);
        const hsize_t three[1] = {3};
        Id vec(h.Screate_simple(1, three, nullptr), h.Sclose);
        attr("bbox_lo", h.f64, m.bbox_lo.data(), vec);
        attr("bbox_hi", h.f64, m.bbox_hi.data(), vec);
    }
    {  // partition tables of one partition
        Id g = group_create(file, "partition");
        const auto table = [&](const char* name, std::uint64_t n) {
            const std::uint64_t v[2] = {0, n};
            write_1d(g, name, h.u64, v, 2);
        };
        table("cell_offsets", nc);
        table("cell_nodes_offsets", m.cell_nodes.size());
        table("face_offsets", nf);
        table("face_nodes_offsets", m.face_nodes.size());
        table("node_offsets", static_cast<std::uint64_t>(m.n_nodes()));
        table("comm_nb_offsets", m.nb_ranks.size());
        const int zero = 0;
        write_1d(g, "part2rank", h.native_int, &zero, 1);
        write_1d(g, "rank2part", h.native_int, &zero, 1);
    }
    {  // cells
        Id g = group_create(file, "cells");
        write_1d(g, "type", h.u8, m.cell_type.data(), nc);
        write_gid(g, "gid", m.cell_gid);
        write_1d(g, "donor", h.i32, m.cell_donor.data(), nc);
        write_1d(g, "volume", h.f64, m.cell_volume.data(), nc);
        write_1d(g, "nodes_offsets", h.i32, m.cell_nodes_off.data(), nc);  // without the trailing offset
        write_1d(g, "nodes", h.i32, m.cell_nodes.data(), m.cell_nodes.size());
        write_vec3(g, "centroid", m.cell_cx, m.cell_cy, m.cell_cz);
    }
    {  // faces
        Id g = group_create(file, "faces");
        write_1d(g, "owner", h.i32, m.face_owner.data(), nf);
        write_1d(g, "neigh", h.i32, m.face_neigh.data(), nf);
        write_1d(g, "patch", h.i32, m.face_patch.data(), nf);
        write_1d(g, "type", h.u8, m.face_type.data(), nf);
        write_1d(g, "area", h.f64, m.face_area.data(), nf);
        write_1d(g, "nodes_offsets", h.i32, m.face_nodes_off.data(), nf);
        write_1d(g, "nodes", h.i32, m.face_nodes.data(), m.face_nodes.size());
        write_vec3(g, "normal", m.face_nx, m.face_ny, m.face_nz);
        write_vec3(g, "centroid", m.face_cx, m.face_cy, m.face_cz);
    }
    {  // nodes
        Id g = group_create(file, "nodes");
        write_gid(g, "gid", m.node_gid);
        write_vec3(g, "coords", m.node_x, m.node_y, m.node_z);
    }
    {  // comm
        Id g = group_create(file, "comm");
        write_1d(g, "nb_ranks", h.i32, m.nb_ranks.data(), m.nb_ranks.size());
        write_1d(g, "send_owned_cells", h.i32, m.send_owned.data(), m.send_owned.size());
        write_1d(g, "recv_ghost_cells", h.i32, m.recv_ghost.data(), m.recv_ghost.size());
        write_1d(g, "send_offsets", h.i32, m.send_offsets.data(), m.send_offsets.size());
        write_1d(g, "recv_offsets", h.i32, m.recv_offsets.data(), m.recv_offsets.size());
    }
    {  // patches
        Id g = group_create(file, "patches");
        write_strings(g, "names", m.patch_names);
        write_strings(g, "cgns_types", m.patch_cgns_types);
        write_1d(g, "patch_face_offsets", h.i32, m.patch_face_offsets.data(), m.patch_face_offsets.size());
        write_1d(g, "patch_faces", h.i32, m.patch_faces.data(), m.patch_faces.size());
    }
}

}  // namespace cfdg::front
