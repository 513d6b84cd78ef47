// Mesh preprocessing on the device: the cell connectivity a context is built
// from (aux_connectivity.cpp:11-187, 348-445; bsr_helpers.hpp:16-59).
//
// Every list is a set of 64-bit keys (row << 32 | entry) sorted by a device
// radix sort, so each row comes out in ascending entry order, which is the
// order the host builders (setup.cpp) produce: cell faces in ascending face
// index, face and vertex neighbours ascending (the vertex stencil without
// duplicates and without the cell itself).  Row offsets are lower bounds of
// (row << 32) in the sorted keys.  The BSR rows are the face-neighbour rows
// with the diagonal inserted in its sorted place.  The vertex stencil is
// formed in chunks of cells whose candidate lists (every cell of every node
// of the cell) hold at most kChunkKeys keys, so its memory stays bounded on
// any mesh size.  Integer work only; every output element has one writer.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "setup.hpp"

namespace cfdg {

namespace {

using u64 = unsigned long long;
constexpr int kBlock = 256;
constexpr long long kChunkKeys = 1LL << 27;  // 1 GiB of keys per vertex-stencil chunk

void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) throw DeviceError(std::string("device connectivity: ") + what + ": " + hipGetErrorString(e));
}

// Device buffer freed on scope exit.
template <class T>
struct Buf {
    T* p = nullptr;
    explicit Buf(std::size_t n) {
        if (n > 0) hip_ok(hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T)), "hipMalloc");
    }
    explicit Buf(const std::vector<T>& h) : Buf(h.size()) {
        if (!h.empty()) hip_ok(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "copy to device");
    }
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    void get(T* h, std::size_t n) const {
        if (n > 0) hip_ok(hipMemcpy(h, p, n * sizeof(T), hipMemcpyDeviceToHost), "copy to host");
    }
};

unsigned grid(long long n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

// bits needed to hold the values 0 .. n
int bit_width(long long n) {
    int b = 0;
    while (b < 62 && (1LL << b) <= n) ++b;
    return b;
}

// owner keys of every face, then neighbour keys (a key of row nc, which sorts
// after every real row, for a face without a neighbour)
__global__ void __launch_bounds__(kBlock) k_face_keys(int nf, int nc, const int* owner, const int* neigh, u64* keys) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    keys[f] = (static_cast<u64>(owner[f]) << 32) | static_cast<unsigned>(f);
    const int n = neigh[f];
    keys[nf + f] = (n >= 0) ? ((static_cast<u64>(n) << 32) | static_cast<unsigned>(f)) : (static_cast<u64>(nc) << 32);
}

// both directions of every interior face
__global__ void __launch_bounds__(kBlock) k_pair_keys(int ninner, const int* owner, const int* neigh, u64* keys) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= ninner) return;
    const u64 o = static_cast<unsigned>(owner[f]), n = static_cast<unsigned>(neigh[f]);
    keys[2 * static_cast<long long>(f)] = (o << 32) | n;
    keys[2 * static_cast<long long>(f) + 1] = (n << 32) | o;
}

// (node, cell) of every entry of the cell node lists
__global__ void __launch_bounds__(kBlock) k_node_keys(int nc, const int* coff, const int* cn, u64* keys) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    for (int k = coff[c]; k < coff[c + 1]; ++k)
        keys[k] = (static_cast<u64>(static_cast<unsigned>(cn[k])) << 32) | static_cast<unsigned>(c);
}

// off[r] = first key of row r or later (rows 0 .. nrows), by binary search
__global__ void __launch_bounds__(kBlock) k_row_offsets(long long nkeys, const u64* keys, int nrows, int* off) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nrows) return;
    const u64 target = static_cast<u64>(r) << 32;
    long long lo = 0, hi = nkeys;
    while (lo < hi) {
        const long long mid = lo + ((hi - lo) >> 1);
        if (keys[mid] < target)
            lo = mid + 1;
        else
            hi = mid;
    }
    off[r] = static_cast<int>(lo);
}

__global__ void __launch_bounds__(kBlock) k_key_entries(long long n, const u64* keys, int* idx) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    idx[i] = static_cast<int>(static_cast<unsigned>(keys[i] & 0xffffffffULL));
}

// BSR row i: its face neighbours with i inserted before the first larger one
__global__ void __launch_bounds__(kBlock) k_bsr_rows(int nc, const int* ccf_off, const int* ccf, int* row_ptr, int* cols,
                                                     int* diag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const int a = ccf_off[i], b = ccf_off[i + 1];
    int w = a + i;
    row_ptr[i] = w;
    if (i == nc - 1) row_ptr[nc] = b + nc;
    bool placed = false;
    for (int e = a; e < b; ++e) {
        const int col = ccf[e];
        if (!placed && col > i) {
            diag[i] = w;
            cols[w++] = i;
            placed = true;
        }
        cols[w++] = col;
    }
    if (!placed) {
        diag[i] = w;
        cols[w] = i;
    }
}

// candidates of cell c's vertex stencil: the cells of its nodes other than c
__global__ void __launch_bounds__(kBlock) k_stencil_count(int nc, const int* coff, const int* cn, const int* noff,
                                                          const int* ncell, int* cnt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    int n = 0;
    for (int k = coff[c]; k < coff[c + 1]; ++k) {
        const int v = cn[k];
        for (int j = noff[v]; j < noff[v + 1]; ++j) n += (ncell[j] != c);
    }
    cnt[c] = n;
}

// the candidates of cells lo .. lo + n - 1 as (cell - lo, candidate) keys at
// the chunk-relative positions pos
__global__ void __launch_bounds__(kBlock) k_stencil_keys(int lo, int n, const long long* pos, const int* coff,
                                                         const int* cn, const int* noff, const int* ncell, u64* keys) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int c = lo + r;
    long long p = pos[r];
    const u64 hi = static_cast<u64>(r) << 32;
    for (int k = coff[c]; k < coff[c + 1]; ++k) {
        const int v = cn[k];
        for (int j = noff[v]; j < noff[v + 1]; ++j) {
            const int d = ncell[j];
            if (d != c) keys[p++] = hi | static_cast<unsigned>(d);
        }
    }
}

// keys sorted in place over bits [0, end_bit) (alt: scratch of the same size)
void sort_keys(Buf<u64>& keys, Buf<u64>& alt, long long n, int end_bit) {
    if (n <= 1) return;
    hipcub::DoubleBuffer<u64> db(keys.p, alt.p);
    std::size_t bytes = 0;
    hip_ok(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, db, n, 0, end_bit), "radix sort size");
    Buf<unsigned char> tmp(bytes);
    hip_ok(hipcub::DeviceRadixSort::SortKeys(tmp.p, bytes, db, n, 0, end_bit), "radix sort");
    if (db.Current() != keys.p) std::swap(keys.p, alt.p);
}

// Csr of sorted keys: offsets of rows 0 .. nrows - 1 and the first nkeys entries
Csr csr_of(const Buf<u64>& keys, long long nkeys, int nrows) {
    Csr out;
    Buf<int> off(static_cast<std::size_t>(nrows) + 1), idx(static_cast<std::size_t>(nkeys));
    k_row_offsets<<<grid(nrows + 1LL), kBlock>>>(nkeys, keys.p, nrows, off.p);
    k_key_entries<<<grid(nkeys), kBlock>>>(nkeys, keys.p, idx.p);
    hip_ok(hipGetLastError(), "launch");
    out.off.resize(static_cast<std::size_t>(nrows) + 1);
    out.idx.resize(static_cast<std::size_t>(nkeys));
    off.get(out.off.data(), out.off.size());
    idx.get(out.idx.data(), out.idx.size());
    return out;
}


// --- reverse Cuthill-McKee order (rcm_order, mesh.cpp; reorder.cpp:88-188) ---

// Dual-graph keys: (cell, face) for both cells of every face whose other
// cell is a cell of the mesh, (nc, 0) otherwise (sorted after every row).
__global__ void __launch_bounds__(kBlock) k_dual_keys(int nf, int nc, const int* owner, const int* neigh, u64* keys) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const int v = neigh[f];
    const bool ok = v >= 0 && v < nc;
    keys[2LL * f] = ok ? ((static_cast<u64>(owner[f]) << 32) | static_cast<unsigned>(f)) : (static_cast<u64>(nc) << 32);
    keys[2LL * f + 1] = ok ? ((static_cast<u64>(v) << 32) | static_cast<unsigned>(f)) : (static_cast<u64>(nc) << 32);
}

// adjacency entry = the other cell of the entry's face
__global__ void __launch_bounds__(kBlock) k_dual_adj(long long n, const u64* keys, const int* owner, const int* neigh,
                                                     int* adj) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = static_cast<int>(keys[i] >> 32), f = static_cast<int>(keys[i] & 0xffffffffULL);
    adj[i] = (owner[f] == c) ? neigh[f] : owner[f];
}

// Every unmarked neighbour v of the frontier queue[q0, q1) takes the smallest
// key (queue position of the parent << 32 | adjacency slot) that reaches it:
// the first parent in queue order and, within it, the first slot.
__global__ void __launch_bounds__(kBlock) k_bfs_claim(int q0, int q1, const int* queue, const int* off, const int* adj,
                                                      const unsigned char* mark, u64* claim) {
    const int p = q0 + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= q1) return;
    const int u = queue[p];
    for (int e = off[u]; e < off[u + 1]; ++e) {
        const int v = adj[e];
        if (!mark[v]) atomicMin(&claim[v], (static_cast<u64>(p) << 32) | static_cast<unsigned>(e - off[u]));
    }
}

// the winning slot of each claimed neighbour appends it (one winner per cell)
__global__ void __launch_bounds__(kBlock) k_bfs_collect(int q0, int q1, const int* queue, const int* off, const int* adj,
                                                        const unsigned char* mark, const u64* claim, u64* ckey,
                                                        int* cval, int* count) {
    const int p = q0 + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= q1) return;
    const int u = queue[p];
    for (int e = off[u]; e < off[u + 1]; ++e) {
        const int v = adj[e];
        const u64 key = (static_cast<u64>(p) << 32) | static_cast<unsigned>(e - off[u]);
        if (!mark[v] && claim[v] == key) {
            const int i = atomicAdd(count, 1);
            ckey[i] = key;
            cval[i] = v;
        }
    }
}

// keys (degree << 32 | cell) of the collected cells, the order among the
// children of one parent in the Cuthill-McKee pass
__global__ void __launch_bounds__(kBlock) k_deg_keys(int n, const int* cval, const int* off, u64* ckey) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = cval[i];
    ckey[i] = (static_cast<u64>(static_cast<unsigned>(off[v + 1] - off[v])) << 32) | static_cast<unsigned>(v);
}

// keys = parent position of the collected cells
__global__ void __launch_bounds__(kBlock) k_parent_keys(int n, const int* cval, const u64* claim, u64* ckey) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ckey[i] = claim[cval[i]] >> 32;
}

// the sorted level goes to queue[q1, q1 + n), marked, its claims cleared
__global__ void __launch_bounds__(kBlock) k_bfs_commit(int n, const int* cval, int* queue_out, unsigned char* mark,
                                                       u64* claim) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = cval[i];
    queue_out[i] = v;
    mark[v] = 1;
    claim[v] = ~0ULL;
}

__global__ void __launch_bounds__(kBlock) k_first_unmarked(int n, const unsigned char* mark, int* first) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !mark[i]) atomicMin(first, i);
}

__global__ void __launch_bounds__(kBlock) k_fill_u64(long long n, u64 v, u64* a) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void __launch_bounds__(kBlock) k_set_one(int v, int pos, int* queue, unsigned char* mark) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        queue[pos] = v;
        mark[v] = 1;
    }
}

// Device scratch that grows on demand and is reused across calls.
struct Scratch {
    void* p = nullptr;
    std::size_t n = 0;
    Scratch() = default;
    Scratch(const Scratch&) = delete;
    Scratch& operator=(const Scratch&) = delete;
    ~Scratch() {
        if (p) (void)hipFree(p);
    }
    void* get(std::size_t want) {
        if (want > n) {
            if (p) (void)hipFree(p);
            p = nullptr;
            n = 0;
            hip_ok(hipMalloc(&p, want), "hipMalloc");
            n = want;
        }
        return p;
    }
};

// pairs sorted by key over bits [0, end_bit), stable
void sort_pairs(Buf<u64>& k, Buf<u64>& k_alt, Buf<int>& v, Buf<int>& v_alt, long long n, int end_bit, Scratch& tmp) {
    if (n <= 1) return;
    hipcub::DoubleBuffer<u64> dk(k.p, k_alt.p);
    hipcub::DoubleBuffer<int> dv(v.p, v_alt.p);
    std::size_t bytes = 0;
    hip_ok(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, dk, dv, n, 0, end_bit), "radix sort size");
    hip_ok(hipcub::DeviceRadixSort::SortPairs(tmp.get(bytes), bytes, dk, dv, n, 0, end_bit), "radix sort");
    if (dk.Current() != k.p) std::swap(k.p, k_alt.p);
    if (dv.Current() != v.p) std::swap(v.p, v_alt.p);
}

}  // namespace

MeshConnectivity device_connectivity(const Mesh& m, bool node_stencil) {
    const int nc = m.n_cells, nf = m.n_faces, ninner = m.n_inner_faces;
    if (nc <= 0 || nf <= 0 || ninner < 0 || ninner > nf) throw std::runtime_error("device connectivity: mesh sizes");
    MeshConnectivity out;
    out.has_node_stencil = node_stencil;
    Buf<int> owner(m.face_owner), neigh(m.face_neigh);
    const int hb_cells = bit_width(nc);  // rows 0 .. nc (nc marks "no neighbour")

    {  // faces of each cell
        const long long n2 = 2LL * nf;
        Buf<u64> keys(static_cast<std::size_t>(n2)), alt(static_cast<std::size_t>(n2));
        k_face_keys<<<grid(nf), kBlock>>>(nf, nc, owner.p, neigh.p, keys.p);
        hip_ok(hipGetLastError(), "launch");
        sort_keys(keys, alt, n2, 32 + hb_cells);
        const long long nvalid = nf + std::count_if(m.face_neigh.begin(), m.face_neigh.end(), [](int v) { return v >= 0; });
        out.cell_faces = csr_of(keys, nvalid, nc);
    }
    Buf<int> ccf_off(static_cast<std::size_t>(nc) + 1), ccf(static_cast<std::size_t>(2LL * ninner));
    {  // face neighbours of each cell
        const long long n2 = 2LL * ninner;
        Buf<u64> keys(static_cast<std::size_t>(n2)), alt(static_cast<std::size_t>(n2));
        if (ninner > 0) {
            k_pair_keys<<<grid(ninner), kBlock>>>(ninner, owner.p, neigh.p, keys.p);
            hip_ok(hipGetLastError(), "launch");
        }
        sort_keys(keys, alt, n2, 32 + hb_cells);
        k_row_offsets<<<grid(nc + 1LL), kBlock>>>(n2, keys.p, nc, ccf_off.p);
        if (n2 > 0) k_key_entries<<<grid(n2), kBlock>>>(n2, keys.p, ccf.p);
        hip_ok(hipGetLastError(), "launch");
        out.cell_cells_face.off.resize(static_cast<std::size_t>(nc) + 1);
        out.cell_cells_face.idx.resize(static_cast<std::size_t>(n2));
        ccf_off.get(out.cell_cells_face.off.data(), out.cell_cells_face.off.size());
        ccf.get(out.cell_cells_face.idx.data(), out.cell_cells_face.idx.size());
    }
    {  // BSR pattern
        const long long nnzb = 2LL * ninner + nc;
        Buf<int> rp(static_cast<std::size_t>(nc) + 1), cols(static_cast<std::size_t>(nnzb)), dg(static_cast<std::size_t>(nc));
        k_bsr_rows<<<grid(nc), kBlock>>>(nc, ccf_off.p, ccf.p, rp.p, cols.p, dg.p);
        hip_ok(hipGetLastError(), "launch");
        out.bsr.row_ptr.resize(static_cast<std::size_t>(nc) + 1);
        out.bsr.cols.resize(static_cast<std::size_t>(nnzb));
        out.bsr.diag.resize(static_cast<std::size_t>(nc));
        rp.get(out.bsr.row_ptr.data(), out.bsr.row_ptr.size());
        cols.get(out.bsr.cols.data(), out.bsr.cols.size());
        dg.get(out.bsr.diag.data(), out.bsr.diag.size());
        for (int i = 0; i < nc; ++i)
            for (int k = out.bsr.row_ptr[static_cast<std::size_t>(i)] + 1; k < out.bsr.row_ptr[static_cast<std::size_t>(i) + 1]; ++k)
                if (out.bsr.cols[static_cast<std::size_t>(k)] == out.bsr.cols[static_cast<std::size_t>(k) - 1])
                    throw std::runtime_error("bsr: two faces couple the same pair of cells");
    }
    if (!node_stencil) return out;

    // vertex stencil: node -> cells, then per chunk of cells the sorted,
    // de-duplicated candidate keys
    const long long ne = m.cell_nodes_off[static_cast<std::size_t>(nc)];
    const int nn = m.n_nodes;
    Buf<int> coff(m.cell_nodes_off), cn(m.cell_nodes);
    Buf<int> noff(static_cast<std::size_t>(nn) + 1), ncell(static_cast<std::size_t>(ne));
    {
        Buf<u64> keys(static_cast<std::size_t>(ne)), alt(static_cast<std::size_t>(ne));
        k_node_keys<<<grid(nc), kBlock>>>(nc, coff.p, cn.p, keys.p);
        hip_ok(hipGetLastError(), "launch");
        sort_keys(keys, alt, ne, 32 + bit_width(nn));
        k_row_offsets<<<grid(nn + 1LL), kBlock>>>(ne, keys.p, nn, noff.p);
        k_key_entries<<<grid(ne), kBlock>>>(ne, keys.p, ncell.p);
        hip_ok(hipGetLastError(), "launch");
    }
    std::vector<int> cnt(static_cast<std::size_t>(nc));
    {
        Buf<int> dcnt(static_cast<std::size_t>(nc));
        k_stencil_count<<<grid(nc), kBlock>>>(nc, coff.p, cn.p, noff.p, ncell.p, dcnt.p);
        hip_ok(hipGetLastError(), "launch");
        dcnt.get(cnt.data(), cnt.size());
    }
    long long maxrow = 0, all = 0;
    for (const int v : cnt) {
        maxrow = std::max<long long>(maxrow, v);
        all += v;
    }
    const long long cap = std::max<long long>(1, std::min(std::max(kChunkKeys, maxrow), all));
    Csr& st = out.cell_cells_node;
    st.off.assign(static_cast<std::size_t>(nc) + 1, 0);
    st.idx.reserve(static_cast<std::size_t>(std::min<long long>(all, 0x7fffffffLL)));  // chunks append without copies
    Buf<u64> keys(static_cast<std::size_t>(cap)), alt(static_cast<std::size_t>(cap)), uniq(static_cast<std::size_t>(cap));
    Buf<long long> dpos(static_cast<std::size_t>(nc)), dnum(1);
    Buf<int> loff(static_cast<std::size_t>(nc) + 1), lidx(static_cast<std::size_t>(cap));
    std::vector<long long> pos;
    for (int lo = 0; lo < nc;) {
        // the largest run of cells from lo whose candidates fit the chunk
        int hi = lo;
        long long total = 0;
        pos.clear();
        while (hi < nc && total + cnt[static_cast<std::size_t>(hi)] <= cap) {
            pos.push_back(total);
            total += cnt[static_cast<std::size_t>(hi)];
            ++hi;
        }
        const int n = hi - lo;
        hip_ok(hipMemcpy(dpos.p, pos.data(), pos.size() * sizeof(long long), hipMemcpyHostToDevice), "copy to device");
        k_stencil_keys<<<grid(n), kBlock>>>(lo, n, dpos.p, coff.p, cn.p, noff.p, ncell.p, keys.p);
        hip_ok(hipGetLastError(), "launch");
        sort_keys(keys, alt, total, 32 + bit_width(n));
        long long nu = 0;
        if (total > 0) {
            std::size_t bytes = 0;
            hip_ok(hipcub::DeviceSelect::Unique(nullptr, bytes, keys.p, uniq.p, dnum.p, total), "unique size");
            Buf<unsigned char> tmp(bytes);
            hip_ok(hipcub::DeviceSelect::Unique(tmp.p, bytes, keys.p, uniq.p, dnum.p, total), "unique");
            dnum.get(&nu, 1);
        }
        k_row_offsets<<<grid(n + 1LL), kBlock>>>(nu, uniq.p, n, loff.p);
        if (nu > 0) k_key_entries<<<grid(nu), kBlock>>>(nu, uniq.p, lidx.p);
        hip_ok(hipGetLastError(), "launch");
        const long long base = static_cast<long long>(st.idx.size());
        if (base + nu > 0x7fffffffLL) throw std::runtime_error("device connectivity: vertex stencil exceeds 2^31 entries");
        std::vector<int> lo_off(static_cast<std::size_t>(n) + 1);
        loff.get(lo_off.data(), lo_off.size());
        st.idx.resize(static_cast<std::size_t>(base + nu));
        lidx.get(st.idx.data() + base, static_cast<std::size_t>(nu));
        for (int r = 0; r < n; ++r) st.off[static_cast<std::size_t>(lo + r) + 1] = static_cast<int>(base + lo_off[static_cast<std::size_t>(r) + 1]);
        lo = hi;
    }
    return out;
}

std::vector<std::int32_t> rcm_order_device(const Mesh& m) {
    const int n = m.n_cells, nf = m.n_faces;
    if (n <= 0 || nf <= 0) throw std::runtime_error("device rcm: mesh sizes");
    // dual graph: each cell's neighbours in ascending face order, as the host builds it
    Buf<int> owner(m.face_owner), neigh(m.face_neigh);
    Buf<int> off(static_cast<std::size_t>(n) + 1);
    long long ne = 0;
    for (int f = 0; f < nf; ++f) {
        const int v = m.face_neigh[static_cast<std::size_t>(f)];
        if (v >= 0 && v < n) ne += 2;
    }
    Buf<int> adj(static_cast<std::size_t>(std::max<long long>(ne, 1)));
    {
        Buf<u64> keys(2 * static_cast<std::size_t>(nf)), alt(2 * static_cast<std::size_t>(nf));
        k_dual_keys<<<grid(nf), kBlock>>>(nf, n, owner.p, neigh.p, keys.p);
        hip_ok(hipGetLastError(), "launch");
        sort_keys(keys, alt, 2LL * nf, 32 + bit_width(n));
        k_row_offsets<<<grid(n + 1LL), kBlock>>>(ne, keys.p, n, off.p);
        if (ne > 0) k_dual_adj<<<grid(ne), kBlock>>>(ne, keys.p, owner.p, neigh.p, adj.p);
        hip_ok(hipGetLastError(), "launch");
    }
    std::vector<int> h_off(static_cast<std::size_t>(n) + 1);
    off.get(h_off.data(), h_off.size());
    const auto deg = [&](int v) { return h_off[static_cast<std::size_t>(v) + 1] - h_off[static_cast<std::size_t>(v)]; };

    Buf<unsigned char> visited(static_cast<std::size_t>(n)), mark(static_cast<std::size_t>(n));
    hip_ok(hipMemset(visited.p, 0, static_cast<std::size_t>(n)), "memset");
    Buf<u64> claim(static_cast<std::size_t>(n));
    k_fill_u64<<<grid(n), kBlock>>>(n, ~0ULL, claim.p);
    Buf<int> pq(static_cast<std::size_t>(n)), order(static_cast<std::size_t>(n)), cnt(1), first(1);
    Buf<u64> ckey(static_cast<std::size_t>(n)), ckey_alt(static_cast<std::size_t>(n));
    Buf<int> cval(static_cast<std::size_t>(n)), cval_alt(static_cast<std::size_t>(n));
    Scratch tmp;
    hip_ok(hipGetLastError(), "launch");

    // one BFS level: the frontier queue[q0, q1) appends its unmarked
    // neighbours at q1 in the order of the mode (0: parent, then adjacency
    // slot, as a plain queue; 1: parent, then degree, then index); returns
    // how many
    const auto level = [&](int* queue, int q0, int q1, int mode) -> int {
        hip_ok(hipMemset(cnt.p, 0, sizeof(int)), "memset");
        k_bfs_claim<<<grid(q1 - q0), kBlock>>>(q0, q1, queue, off.p, adj.p, mark.p, claim.p);
        k_bfs_collect<<<grid(q1 - q0), kBlock>>>(q0, q1, queue, off.p, adj.p, mark.p, claim.p, ckey.p, cval.p, cnt.p);
        hip_ok(hipGetLastError(), "launch");
        int c = 0;
        cnt.get(&c, 1);
        if (c < 0 || static_cast<long long>(q1) + c > n) throw DeviceError("device rcm: level exceeds the cell count");
        if (c == 0) return 0;
        const int pbits = bit_width(q1);
        if (mode == 0) {
            sort_pairs(ckey, ckey_alt, cval, cval_alt, c, 32 + pbits, tmp);
        } else {
            k_deg_keys<<<grid(c), kBlock>>>(c, cval.p, off.p, ckey.p);
            hip_ok(hipGetLastError(), "launch");
            sort_pairs(ckey, ckey_alt, cval, cval_alt, c, 64, tmp);
            k_parent_keys<<<grid(c), kBlock>>>(c, cval.p, claim.p, ckey.p);
            hip_ok(hipGetLastError(), "launch");
            sort_pairs(ckey, ckey_alt, cval, cval_alt, c, pbits, tmp);
        }
        k_bfs_commit<<<grid(c), kBlock>>>(c, cval.p, queue + q1, mark.p, claim.p);
        hip_ok(hipGetLastError(), "launch");
        return c;
    };

    // pseudo-peripheral start: up to three BFS passes from the deepest cell of
    // the previous one (lowest degree among the deepest, first in queue order)
    const auto peripheral = [&](int root) {
        int best = root, max_depth = 0;
        std::vector<int> last;
        for (int pass = 0; pass < 3; ++pass) {
            hip_ok(hipMemcpy(mark.p, visited.p, static_cast<std::size_t>(n), hipMemcpyDeviceToDevice), "copy");
            k_set_one<<<1, 1>>>(best, 0, pq.p, mark.p);
            hip_ok(hipGetLastError(), "launch");
            int q0 = 0, q1 = 1, depth = 0;
            for (;;) {
                const int c = level(pq.p, q0, q1, 0);
                if (c == 0) break;
                q0 = q1;
                q1 += c;
                ++depth;
            }
            int far = best;
            if (depth > 0) {
                last.resize(static_cast<std::size_t>(q1 - q0));
                hip_ok(hipMemcpy(last.data(), pq.p + q0, last.size() * sizeof(int), hipMemcpyDeviceToHost), "copy");
                far = last[0];
                for (const int v : last)
                    if (deg(v) < deg(far)) far = v;
            }
            if (depth <= max_depth) break;
            max_depth = depth;
            best = far;
        }
        return best;
    };

    int total = 0;
    while (total < n) {
        int f = n;
        hip_ok(hipMemcpy(first.p, &f, sizeof(int), hipMemcpyHostToDevice), "copy");
        k_first_unmarked<<<grid(n), kBlock>>>(n, visited.p, first.p);
        hip_ok(hipGetLastError(), "launch");
        first.get(&f, 1);
        if (f >= n) throw DeviceError("device rcm: no unvisited cell left before the order is complete");
        const int start = peripheral(f);
        // Cuthill-McKee pass of this component, marks kept in `visited`
        hip_ok(hipMemcpy(mark.p, visited.p, static_cast<std::size_t>(n), hipMemcpyDeviceToDevice), "copy");
        k_set_one<<<1, 1>>>(start, total, order.p, mark.p);
        hip_ok(hipGetLastError(), "launch");
        int q0 = total, q1 = total + 1;
        for (;;) {
            const int c = level(order.p, q0, q1, 1);
            if (c == 0) break;
            q0 = q1;
            q1 += c;
        }
        hip_ok(hipMemcpy(visited.p, mark.p, static_cast<std::size_t>(n), hipMemcpyDeviceToDevice), "copy");
        total = q1;
    }
    std::vector<std::int32_t> out(static_cast<std::size_t>(n));
    order.get(out.data(), out.size());
    std::reverse(out.begin(), out.end());
    return out;
}

}  // namespace cfdg
