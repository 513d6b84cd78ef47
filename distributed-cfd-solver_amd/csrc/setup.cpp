#include "setup.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <stdexcept>
#include <thread>

#include "parallel.hpp"

namespace cfdg {

namespace {
using i32 = std::int32_t;
inline std::size_t z(i32 v) { return static_cast<std::size_t>(v); }
}  // namespace

Csr cell_faces(const Mesh& m) {
    Csr c;
    c.off.assign(z(m.n_cells) + 1, 0);
    for (i32 f = 0; f < m.n_faces; ++f) {
        ++c.off[z(m.face_owner[z(f)]) + 1];
        if (m.face_neigh[z(f)] >= 0) ++c.off[z(m.face_neigh[z(f)]) + 1];
    }
    for (i32 i = 0; i < m.n_cells; ++i) c.off[z(i) + 1] += c.off[z(i)];
    c.idx.resize(z(c.off.back()));
    std::vector<i32> cur(c.off.begin(), c.off.end() - 1);
    for (i32 f = 0; f < m.n_faces; ++f) {
        c.idx[z(cur[z(m.face_owner[z(f)])]++)] = f;
        if (m.face_neigh[z(f)] >= 0) c.idx[z(cur[z(m.face_neigh[z(f)])]++)] = f;
    }
    return c;
}

Csr cell_cells_by_face(const Mesh& m) {
    Csr c;
    c.off.assign(z(m.n_cells) + 1, 0);
    for (i32 f = 0; f < m.n_inner_faces; ++f) {
        ++c.off[z(m.face_owner[z(f)]) + 1];
        ++c.off[z(m.face_neigh[z(f)]) + 1];
    }
    for (i32 i = 0; i < m.n_cells; ++i) c.off[z(i) + 1] += c.off[z(i)];
    c.idx.resize(z(c.off.back()));
    std::vector<i32> cur(c.off.begin(), c.off.end() - 1);
    for (i32 f = 0; f < m.n_inner_faces; ++f) {
        const i32 o = m.face_owner[z(f)], n = m.face_neigh[z(f)];
        c.idx[z(cur[z(o)]++)] = n;
        c.idx[z(cur[z(n)]++)] = o;
    }
    for (i32 i = 0; i < m.n_cells; ++i) std::sort(c.idx.begin() + c.off[z(i)], c.idx.begin() + c.off[z(i) + 1]);
    return c;
}

Csr cell_cells_by_node(const Mesh& m) {
    Csr nc;
    nc.off.assign(z(m.n_nodes) + 1, 0);
    for (i32 c = 0; c < m.n_cells; ++c)
        for (i32 k = m.cell_nodes_off[z(c)]; k < m.cell_nodes_off[z(c) + 1]; ++k) ++nc.off[z(m.cell_nodes[z(k)]) + 1];
    for (i32 n = 0; n < m.n_nodes; ++n) nc.off[z(n) + 1] += nc.off[z(n)];
    nc.idx.resize(z(nc.off.back()));
    {
        std::vector<i32> cur(nc.off.begin(), nc.off.end() - 1);
        for (i32 c = 0; c < m.n_cells; ++c)
            for (i32 k = m.cell_nodes_off[z(c)]; k < m.cell_nodes_off[z(c) + 1]; ++k)
                nc.idx[z(cur[z(m.cell_nodes[z(k)])]++)] = c;
    }
    // per cell: the cells of its nodes, sorted, without duplicates and itself
    std::vector<std::vector<i32>> scratch(z(parallel_workers()));
    const auto gather = [&](i32 c, std::vector<i32>& v) {
        v.clear();
        for (i32 k = m.cell_nodes_off[z(c)]; k < m.cell_nodes_off[z(c) + 1]; ++k) {
            const i32 n = m.cell_nodes[z(k)];
            for (i32 j = nc.off[z(n)]; j < nc.off[z(n) + 1]; ++j)
                if (nc.idx[z(j)] != c) v.push_back(nc.idx[z(j)]);
        }
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    };
    Csr out;
    out.off.assign(z(m.n_cells) + 1, 0);
    parallel_for(z(m.n_cells), [&](std::size_t c, int w) {
        gather(static_cast<i32>(c), scratch[z(w)]);
        out.off[c + 1] = static_cast<i32>(scratch[z(w)].size());
    });
    for (i32 c = 0; c < m.n_cells; ++c) out.off[z(c) + 1] += out.off[z(c)];
    out.idx.resize(z(out.off.back()));
    parallel_for(z(m.n_cells), [&](std::size_t c, int w) {
        gather(static_cast<i32>(c), scratch[z(w)]);
        std::copy(scratch[z(w)].begin(), scratch[z(w)].end(), out.idx.begin() + out.off[c]);
    });
    return out;
}

std::vector<double> lsq_coefficients(const Mesh& m, const Csr& st) {
    const std::size_t E = st.idx.size();
    std::vector<double> w(3 * E, 0.0);
    double* wx = w.data();
    double* wy = wx + E;
    double* wz = wy + E;
    std::vector<double> dr, wb;
    for (i32 c = 0; c < m.n_cells; ++c) {
        const std::size_t b = z(st.off[z(c)]), e = z(st.off[z(c) + 1]);
        if (e == b) continue;
        const double x0 = m.cell_cx[z(c)], y0 = m.cell_cy[z(c)], z0 = m.cell_cz[z(c)];
        dr.assign(3 * (e - b), 0.0);
        wb.assign(e - b, 0.0);
        double A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (std::size_t i = 0; i < e - b; ++i) {
            const std::size_t nb = z(st.idx[b + i]);
            const double dx = m.cell_cx[nb] - x0, dy = m.cell_cy[nb] - y0, dz = m.cell_cz[nb] - z0;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double wt = 1.0 / std::max(r2, 1.0e-30);
            dr[3 * i] = dx;
            dr[3 * i + 1] = dy;
            dr[3 * i + 2] = dz;
            wb[i] = wt;
            A[0] += wt * dx * dx;
            A[1] += wt * dx * dy;
            A[2] += wt * dx * dz;
            A[4] += wt * dy * dy;
            A[5] += wt * dy * dz;
            A[8] += wt * dz * dz;
        }
        A[3] = A[1];
        A[6] = A[2];
        A[7] = A[5];
        A[0] += 1.0e-14;
        A[4] += 1.0e-14;
        A[8] += 1.0e-14;
        // adjugate inverse; |det| <= 1e-14 gives a zero operator (gradient_operator.cpp:20-47)
        double inv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                           A[2] * (A[3] * A[7] - A[4] * A[6]);
        if (!(std::abs(det) <= 1.0e-14)) {
            const double id = 1.0 / det;
            inv[0] = (A[4] * A[8] - A[5] * A[7]) * id;
            inv[1] = (A[2] * A[7] - A[1] * A[8]) * id;
            inv[2] = (A[1] * A[5] - A[2] * A[4]) * id;
            inv[3] = (A[5] * A[6] - A[3] * A[8]) * id;
            inv[4] = (A[0] * A[8] - A[2] * A[6]) * id;
            inv[5] = (A[2] * A[3] - A[0] * A[5]) * id;
            inv[6] = (A[3] * A[7] - A[4] * A[6]) * id;
            inv[7] = (A[1] * A[6] - A[0] * A[7]) * id;
            inv[8] = (A[0] * A[4] - A[1] * A[3]) * id;
        }
        for (std::size_t i = 0; i < e - b; ++i) {
            const double rx = wb[i] * dr[3 * i], ry = wb[i] * dr[3 * i + 1], rz = wb[i] * dr[3 * i + 2];
            wx[b + i] = inv[0] * rx + inv[1] * ry + inv[2] * rz;
            wy[b + i] = inv[3] * rx + inv[4] * ry + inv[5] * rz;
            wz[b + i] = inv[6] * rx + inv[7] * ry + inv[8] * rz;
        }
    }
    return w;
}

void face_distances(const Mesh& m, std::vector<double>& dx, std::vector<double>& dy, std::vector<double>& dz,
                    std::vector<double>& dinv) {
    const std::size_t nf = z(m.n_faces);
    dx.resize(nf);
    dy.resize(nf);
    dz.resize(nf);
    dinv.resize(nf);
    for (std::size_t f = 0; f < nf; ++f) {
        const std::size_t u = z(m.face_owner[f]);
        if (f < z(m.n_inner_faces)) {
            const std::size_t v = z(m.face_neigh[f]);
            const double ex = m.cell_cx[v] - m.cell_cx[u], ey = m.cell_cy[v] - m.cell_cy[u], ez = m.cell_cz[v] - m.cell_cz[u];
            dx[f] = ex;
            dy[f] = ey;
            dz[f] = ez;
            dinv[f] = 1.0 / std::sqrt(ex * ex + ey * ey + ez * ez);
        } else {
            const double ex = m.face_cx[f] - m.cell_cx[u], ey = m.face_cy[f] - m.cell_cy[u], ez = m.face_cz[f] - m.cell_cz[u];
            dx[f] = 2.0 * ex;
            dy[f] = 2.0 * ey;
            dz[f] = 2.0 * ez;
            dinv[f] = 1.0 / std::sqrt(4.0 * ex * ex + 4.0 * ey * ey + 4.0 * ez * ez);
        }
    }
}

BsrPattern bsr_pattern(const Csr& ccf, i32 n) {
    BsrPattern p;
    p.row_ptr.assign(z(n) + 1, 0);
    p.diag.resize(z(n));
    for (i32 i = 0; i < n; ++i) p.row_ptr[z(i) + 1] = p.row_ptr[z(i)] + 1 + (ccf.off[z(i) + 1] - ccf.off[z(i)]);
    p.cols.resize(z(p.row_ptr.back()));
    for (i32 i = 0; i < n; ++i) {
        std::size_t w = z(p.row_ptr[z(i)]);
        const std::size_t s = w;
        p.cols[w++] = i;
        for (i32 e = ccf.off[z(i)]; e < ccf.off[z(i) + 1]; ++e) p.cols[w++] = ccf.idx[z(e)];
        std::sort(p.cols.begin() + static_cast<std::ptrdiff_t>(s), p.cols.begin() + static_cast<std::ptrdiff_t>(w));
        for (std::size_t k = s; k < w; ++k) {
            if (k > s && p.cols[k] == p.cols[k - 1])
                throw std::runtime_error("bsr: two faces couple the same pair of cells");
            if (p.cols[k] == i) p.diag[z(i)] = static_cast<i32>(k);
        }
    }
    return p;
}

MeshConnectivity host_connectivity(const Mesh& m, bool node_stencil) {
    MeshConnectivity c;
    c.cell_faces = cell_faces(m);
    c.cell_cells_face = cell_cells_by_face(m);
    c.bsr = bsr_pattern(c.cell_cells_face, m.n_cells);
    if (node_stencil) c.cell_cells_node = cell_cells_by_node(m);
    c.has_node_stencil = node_stencil;
    return c;
}

std::vector<i32> greedy_colours(const Csr& ccf, i32 n, i32& n_colours) {
    std::vector<i32> col(z(n), -1);
    n_colours = 0;
    std::vector<char> used;
    for (i32 i = 0; i < n; ++i) {
        used.assign(z(n_colours) + 1, 0);
        for (i32 e = ccf.off[z(i)]; e < ccf.off[z(i) + 1]; ++e) {
            const i32 j = ccf.idx[z(e)];
            if (j < i) used[z(col[z(j)])] = 1;
        }
        i32 k = 0;
        while (used[z(k)]) ++k;
        col[z(i)] = k;
        n_colours = std::max(n_colours, k + 1);
    }
    return col;
}

RowGroups groups_from_labels(const std::vector<i32>& label, i32 n_labels, bool reverse) {
    RowGroups g;
    g.off.assign(z(n_labels) + 1, 0);
    for (const i32 l : label) {
        const i32 slot = reverse ? (n_labels - 1 - l) : l;
        ++g.off[z(slot) + 1];
    }
    for (i32 k = 0; k < n_labels; ++k) g.off[z(k) + 1] += g.off[z(k)];
    g.rows.resize(label.size());
    std::vector<i32> cur(g.off.begin(), g.off.end() - 1);
    for (std::size_t i = 0; i < label.size(); ++i) {
        const i32 slot = reverse ? (n_labels - 1 - label[i]) : label[i];
        g.rows[z(cur[z(slot)]++)] = static_cast<i32>(i);
    }
    return g;
}

std::vector<i32> forward_levels(const Csr& ccf, i32 n, i32& n_levels) {
    std::vector<i32> lev(z(n), 0);
    n_levels = (n > 0) ? 1 : 0;
    for (i32 i = 0; i < n; ++i) {
        i32 l = 0;
        for (i32 e = ccf.off[z(i)]; e < ccf.off[z(i) + 1]; ++e) {
            const i32 j = ccf.idx[z(e)];
            if (j < i) l = std::max(l, lev[z(j)] + 1);
        }
        lev[z(i)] = l;
        n_levels = std::max(n_levels, l + 1);
    }
    return lev;
}

namespace {

// Faces touching cells[lo, hi), each counted once.
int tile_face_count(const Csr& cf, const Mesh& m, const std::vector<i32>& cells, std::size_t lo, std::size_t hi,
                    std::vector<i32>& stamp, i32 tag) {
    for (std::size_t i = lo; i < hi; ++i) stamp[z(cells[i])] = tag;
    int n = 0;
    for (std::size_t i = lo; i < hi; ++i) {
        const i32 c = cells[i];
        for (i32 e = cf.off[z(c)]; e < cf.off[z(c) + 1]; ++e) {
            const i32 f = cf.idx[z(e)];
            const i32 o = m.face_owner[z(f)], nb = m.face_neigh[z(f)];
            const i32 other = (o == c) ? nb : o;
            // an interior face with both cells in the set is counted at its owner only
            if (other >= 0 && stamp[z(other)] == tag && o != c) continue;
            ++n;
        }
    }
    return n;
}

struct Bisector {
    const Mesh& m;
    const Csr& ccf;
    const Csr& cf;
    int max_cells, max_faces;
    std::vector<i32>& cells;  // shared: every bisector works on its own range
    std::vector<i32> work, best;
    std::vector<i32> stamp;   // per cell: tag of the set it was last marked in
    i32 tag = 0;
    std::vector<i32> tile_off;

    Bisector(const Mesh& mm, const Csr& a, const Csr& b, int mc, int mf, std::vector<i32>& c)
        : m(mm), ccf(a), cf(b), max_cells(mc), max_faces(mf), cells(c), stamp(c.size(), 0) {}

    double coord(i32 c, int a) const {
        return a == 0 ? m.cell_cx[z(c)] : (a == 1 ? m.cell_cy[z(c)] : m.cell_cz[z(c)]);
    }

    // The two halves of a split are independent: down to depth 4 the right
    // half runs on a thread of its own (with its own marks), and its tiles are
    // appended after the left half's, so the tiling is the sequential one.
    void split(std::size_t lo, std::size_t hi, int depth = 0) {
        const std::size_t n = hi - lo;
        const std::size_t cap = static_cast<std::size_t>(max_cells);
        if (n <= cap && tile_face_count(cf, m, cells, lo, hi, stamp, ++tag) <= max_faces) {
            tile_off.push_back(static_cast<i32>(hi));
            return;
        }
        if (n < 2) throw std::runtime_error("tiling: one cell has more faces than a tile holds");
        const std::size_t parts = std::max<std::size_t>(2, (n + cap - 1) / cap);
        const std::size_t nl = n * (parts / 2) / parts;
        long long best_cut = -1;
        for (int a = 0; a < 3; ++a) {
            work.assign(cells.begin() + static_cast<long>(lo), cells.begin() + static_cast<long>(hi));
            std::nth_element(work.begin(), work.begin() + static_cast<long>(nl), work.end(), [&](i32 x, i32 y) {
                const double cx = coord(x, a), cy = coord(y, a);
                return cx < cy || (cx == cy && x < y);
            });
            // faces crossing the split: left cells whose face neighbour is on the right
            const i32 tl = ++tag, tr = ++tag;
            for (std::size_t i = 0; i < nl; ++i) stamp[z(work[i])] = tl;
            for (std::size_t i = nl; i < n; ++i) stamp[z(work[i])] = tr;
            long long cut = 0;
            for (std::size_t i = 0; i < nl; ++i) {
                const i32 c = work[i];
                for (i32 e = ccf.off[z(c)]; e < ccf.off[z(c) + 1]; ++e)
                    if (stamp[z(ccf.idx[z(e)])] == tr) ++cut;
            }
            if (best_cut < 0 || cut < best_cut) {
                best_cut = cut;
                best.swap(work);
            }
        }
        std::copy(best.begin(), best.end(), cells.begin() + static_cast<long>(lo));
        if (depth < 4 && n > (std::size_t{1} << 16)) {
            Bisector right(m, ccf, cf, max_cells, max_faces, cells);
            std::exception_ptr err;
            std::thread th([&] {
                try {
                    right.split(lo + nl, hi, depth + 1);
                } catch (...) {
                    err = std::current_exception();
                }
            });
            try {
                split(lo, lo + nl, depth + 1);
            } catch (...) {
                th.join();
                throw;
            }
            th.join();
            if (err) std::rethrow_exception(err);
            tile_off.insert(tile_off.end(), right.tile_off.begin(), right.tile_off.end());
        } else {
            split(lo, lo + nl, depth + 1);
            split(lo + nl, hi, depth + 1);
        }
    }
};

}  // namespace

Tiling tile_cells(const Mesh& m, const Csr& ccf, const Csr& cf, int max_cells, int max_faces) {
    const i32 n = m.n_cells;
    Tiling t;
    std::vector<i32> cells(z(n));
    for (i32 c = 0; c < n; ++c) cells[z(c)] = c;
    Bisector b(m, ccf, cf, max_cells, max_faces, cells);
    b.tile_off.push_back(0);
    if (n > 0) b.split(0, z(n));
    t.tile_off = std::move(b.tile_off);
    t.new2old = std::move(cells);
    parallel_for(t.tile_off.size() - 1, [&](std::size_t k, int) {
        std::sort(t.new2old.begin() + t.tile_off[k], t.new2old.begin() + t.tile_off[k + 1]);
    });
    t.old2new.assign(z(n), -1);
    for (i32 d = 0; d < n; ++d) t.old2new[z(t.new2old[z(d)])] = d;
    return t;
}

TileFaces tile_faces(const Tiling& t, const Csr& dev_cf, const std::vector<i32>& face_owner_dev,
                     const std::vector<i32>& face_other_dev) {
    TileFaces tf;
    const std::size_t ntiles = t.tile_off.size() - 1;
    tf.tile_face_off.assign(ntiles + 1, 0);
    tf.cell_slot.assign(dev_cf.idx.size(), 0);
    std::vector<i32> pos(face_owner_dev.size(), -1);
    std::vector<std::pair<int, i32>> local;  // (group, face)
    for (std::size_t k = 0; k < ntiles; ++k) {
        const i32 c0 = t.tile_off[k], c1 = t.tile_off[k + 1];
        const auto inside = [&](i32 c) { return c >= c0 && c < c1; };
        local.clear();
        for (i32 c = c0; c < c1; ++c)
            for (i32 e = dev_cf.off[z(c)]; e < dev_cf.off[z(c) + 1]; ++e) {
                const i32 f = dev_cf.idx[z(e)];
                const bool in0 = inside(face_owner_dev[z(f)]), in1 = inside(face_other_dev[z(f)]);
                local.emplace_back(in0 ? (in1 ? 0 : 1) : 2, f);
            }
        std::sort(local.begin(), local.end());
        local.erase(std::unique(local.begin(), local.end()), local.end());
        for (std::size_t j = 0; j < local.size(); ++j) {
            pos[z(local[j].second)] = static_cast<i32>(j);
            tf.faces.push_back(local[j].second);
        }
        for (i32 c = c0; c < c1; ++c) {
            for (i32 e = dev_cf.off[z(c)]; e < dev_cf.off[z(c) + 1]; ++e) {
                const i32 f = dev_cf.idx[z(e)];
                const i32 slot = pos[z(f)];
                tf.cell_slot[z(e)] = (face_owner_dev[z(f)] == c) ? slot : -slot - 1;
            }
        }
        for (std::size_t j = 0; j < local.size(); ++j) pos[z(local[j].second)] = -1;
        tf.tile_face_off[k + 1] = static_cast<i32>(tf.faces.size());
    }
    return tf;
}

namespace {

// Halo of a tile [c0, c1): the cells outside it that its cells' node and face
// stencils reach, ascending.
void tile_halo(i32 c0, i32 c1, const Csr& dev_ccn, const Csr& dev_ccf, std::vector<i32>& local) {
    local.clear();
    for (i32 c = c0; c < c1; ++c) {
        for (i32 e = dev_ccn.off[z(c)]; e < dev_ccn.off[z(c) + 1]; ++e) {
            const i32 nb = dev_ccn.idx[z(e)];
            if (nb < c0 || nb >= c1) local.push_back(nb);
        }
        for (i32 e = dev_ccf.off[z(c)]; e < dev_ccf.off[z(c) + 1]; ++e) {
            const i32 nb = dev_ccf.idx[z(e)];
            if (nb < c0 || nb >= c1) local.push_back(nb);
        }
    }
    std::sort(local.begin(), local.end());
    local.erase(std::unique(local.begin(), local.end()), local.end());
}

}  // namespace

i32 tile_max_slots(const std::vector<i32>& tile_off, const Csr& dev_ccn, const Csr& dev_ccf) {
    const std::size_t ntiles = tile_off.size() - 1;
    std::vector<i32> slots(ntiles, 0);
    std::vector<std::vector<i32>> scratch(z(parallel_workers()));
    parallel_for(ntiles, [&](std::size_t k, int w) {
        tile_halo(tile_off[k], tile_off[k + 1], dev_ccn, dev_ccf, scratch[z(w)]);
        slots[k] = (tile_off[k + 1] - tile_off[k]) + static_cast<i32>(scratch[z(w)].size());
    }, 4);
    i32 most = 0;
    for (const i32 v : slots) most = std::max(most, v);
    return most;
}

TileStencil tile_stencil(const std::vector<i32>& tile_off, const Csr& dev_ccn, const Csr& dev_ccf) {
    TileStencil ts;
    const std::size_t ntiles = tile_off.size() - 1;
    const std::vector<i32>& toff = tile_off;
    ts.halo_off.assign(ntiles + 1, 0);
    ts.ccn_base.assign(ntiles + 1, 0);
    ts.ccf_base.assign(ntiles + 1, 0);
    for (std::size_t k = 0; k < ntiles; ++k) {
        i32 wn = 0, wf = 0;
        for (i32 c = toff[k]; c < toff[k + 1]; ++c) {
            wn = std::max(wn, dev_ccn.off[z(c) + 1] - dev_ccn.off[z(c)]);
            wf = std::max(wf, dev_ccf.off[z(c) + 1] - dev_ccf.off[z(c)]);
        }
        const i32 nt = toff[k + 1] - toff[k];
        ts.ccn_base[k + 1] = ts.ccn_base[k] + (wn + 3) / 4 * 4 * nt;
        ts.ccf_base[k + 1] = ts.ccf_base[k] + (wf + 3) / 4 * 4 * nt;
    }
    ts.ccn_slot.assign(z(ts.ccn_base[ntiles]), 0);
    ts.ccf_slot.assign(z(ts.ccf_base[ntiles]), 0);
    // halo of every tile (independent per tile)
    std::vector<std::vector<i32>> halos(ntiles);
    parallel_for(ntiles, [&](std::size_t k, int) { tile_halo(toff[k], toff[k + 1], dev_ccn, dev_ccf, halos[k]); }, 4);
    for (std::size_t k = 0; k < ntiles; ++k) {
        const i32 nslots = (toff[k + 1] - toff[k]) + static_cast<i32>(halos[k].size());
        if (nslots > 65535) throw std::runtime_error("tile stencil: more than 65535 LDS slots");
        ts.max_slots = std::max(ts.max_slots, nslots);
        ts.halo_off[k + 1] = ts.halo_off[k] + static_cast<i32>(halos[k].size());
    }
    ts.halo.resize(z(ts.halo_off[ntiles]));
    // slot lists: a tile cell is its offset in the tile, a halo cell follows
    // the tile cells in ascending order
    parallel_for(ntiles, [&](std::size_t k, int) {
        const i32 c0 = toff[k], c1 = toff[k + 1], nt = c1 - c0;
        const std::vector<i32>& local = halos[k];
        std::copy(local.begin(), local.end(), ts.halo.begin() + ts.halo_off[k]);
        const auto slot = [&](i32 nb) -> std::uint16_t {
            if (nb >= c0 && nb < c1) return static_cast<std::uint16_t>(nb - c0);
            const auto it = std::lower_bound(local.begin(), local.end(), nb);
            return static_cast<std::uint16_t>(nt + static_cast<i32>(it - local.begin()));
        };
        for (i32 c = c0; c < c1; ++c) {
            const i32 i = c - c0;
            for (i32 e = dev_ccn.off[z(c)]; e < dev_ccn.off[z(c) + 1]; ++e) {
                const i32 kk = e - dev_ccn.off[z(c)];
                ts.ccn_slot[z(ts.ccn_base[k] + ((kk >> 2) * nt + i) * 4 + (kk & 3))] = slot(dev_ccn.idx[z(e)]);
            }
            for (i32 e = dev_ccf.off[z(c)]; e < dev_ccf.off[z(c) + 1]; ++e) {
                const i32 kk = e - dev_ccf.off[z(c)];
                ts.ccf_slot[z(ts.ccf_base[k] + ((kk >> 2) * nt + i) * 4 + (kk & 3))] = slot(dev_ccf.idx[z(e)]);
            }
        }
        std::vector<i32>().swap(halos[k]);
    }, 4);
    return ts;
}

std::vector<i32> split_tiles(const std::vector<i32>& tile_off, const Csr& dev_ccn, const Csr& dev_ccf, int max_slots) {
    const std::size_t n = dev_ccn.off.size() - 1;
    const std::size_t ntiles = tile_off.size() - 1;
    const auto for_nb = [&](i32 c, auto&& fn) {
        for (i32 e = dev_ccn.off[z(c)]; e < dev_ccn.off[z(c) + 1]; ++e) fn(dev_ccn.idx[z(e)]);
        for (i32 e = dev_ccf.off[z(c)]; e < dev_ccf.off[z(c) + 1]; ++e) fn(dev_ccf.idx[z(e)]);
    };
    // per tile, independently: the cells that start a new sub-range.  Each
    // worker marks cells with chunk ids of its own (increasing), so its mark
    // array never needs clearing.
    std::vector<std::vector<i32>> cuts(ntiles);
    const int nw = parallel_workers();
    std::vector<std::vector<i32>> marks(z(nw));
    std::vector<i32> chunks(z(nw), 0);
    parallel_for(ntiles, [&](std::size_t k, int w) {
        std::vector<i32>& mark = marks[z(w)];
        if (mark.empty()) mark.assign(n, -1);  // chunk id that counts the cell as halo
        i32& chunk = chunks[z(w)];
        const i32 t0 = tile_off[k], t1 = tile_off[k + 1];
        i32 start = t0, halo = 0;
        ++chunk;
        for (i32 c = t0; c < t1; ++c) {
            // slots if c joins [start, c]: own cells + halo (c leaves the halo if it was in it)
            i32 fresh = 0;
            for_nb(c, [&](i32 nb) {
                if ((nb < start || nb > c) && mark[z(nb)] != chunk) {
                    mark[z(nb)] = -2 - chunk;  // counted once in this trial
                    ++fresh;
                }
            });
            for_nb(c, [&](i32 nb) {
                if (mark[z(nb)] == -2 - chunk) mark[z(nb)] = -1;
            });
            const i32 was_halo = (mark[z(c)] == chunk) ? 1 : 0;
            if (c > start && (c + 1 - start) + (halo - was_halo + fresh) > max_slots) {
                cuts[k].push_back(c);  // close the chunk before c
                start = c;
                halo = 0;
                ++chunk;
            }
            if (mark[z(c)] == chunk) {
                --halo;
                mark[z(c)] = -1;
            }
            for_nb(c, [&](i32 nb) {
                if ((nb < start || nb > c) && mark[z(nb)] != chunk) {
                    mark[z(nb)] = chunk;
                    ++halo;
                }
            });
        }
    }, 4);
    std::vector<i32> out{0};
    for (std::size_t k = 0; k < ntiles; ++k) {
        out.insert(out.end(), cuts[k].begin(), cuts[k].end());
        out.push_back(tile_off[k + 1]);
    }
    return out;
}

std::vector<i32> backward_levels(const Csr& ccf, i32 n, i32& n_levels) {
    std::vector<i32> lev(z(n), 0);
    n_levels = (n > 0) ? 1 : 0;
    for (i32 i = n - 1; i >= 0; --i) {
        i32 l = 0;
        for (i32 e = ccf.off[z(i)]; e < ccf.off[z(i) + 1]; ++e) {
            const i32 j = ccf.idx[z(e)];
            if (j > i) l = std::max(l, lev[z(j)] + 1);
        }
        lev[z(i)] = l;
        n_levels = std::max(n_levels, l + 1);
    }
    return lev;
}

// Wall distance by graph relaxation (aux_geometry.cpp:545-697, one rank):
// cells next to a wall face start at |x_f - x_c| (minimum over their wall
// faces), all others at 1e30; each sweep relaxes d_c = min(d_c, d_n + |x_n - x_c|)
// over the stencil in ascending and then descending cell order (Gauss-Seidel,
// updated values used at once) until the largest relative decrease of a sweep
// is at most rel_tol or max_sweeps sweeps have run.
std::vector<double> wall_distance(const Mesh& m, const Csr& st, const std::vector<char>& wall_patch, int max_sweeps,
                                  double rel_tol) {
    constexpr double kInf = 1.0e30, kMinDist = 1.0e-12;
    std::vector<double> d(z(m.n_cells), kInf);
    for (i32 f = m.n_inner_faces; f < m.n_faces; ++f) {
        const i32 p = m.face_patch[z(f)];
        if (p < 0 || z(p) >= wall_patch.size() || !wall_patch[z(p)]) continue;
        const std::size_t c = z(m.face_owner[z(f)]);
        const double ex = m.face_cx[z(f)] - m.cell_cx[c], ey = m.face_cy[z(f)] - m.cell_cy[c],
                     ez = m.face_cz[z(f)] - m.cell_cz[c];
        d[c] = std::min(d[c], std::sqrt(ex * ex + ey * ey + ez * ez));
    }
    std::vector<double> len(st.idx.size());
    for (i32 c = 0; c < m.n_cells; ++c)
        for (i32 j = st.off[z(c)]; j < st.off[z(c) + 1]; ++j) {
            const std::size_t n = z(st.idx[z(j)]);
            const double dx = m.cell_cx[n] - m.cell_cx[z(c)], dy = m.cell_cy[n] - m.cell_cy[z(c)],
                         dz = m.cell_cz[n] - m.cell_cz[z(c)];
            len[z(j)] = std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
        double worst = 0.0;
        const auto relax = [&](std::size_t c) {
            const double d0 = d[c];
            double dn = d0;
            for (i32 j = st.off[c]; j < st.off[c + 1]; ++j) {
                const double cand = d[z(st.idx[z(j)])] + len[z(j)];
                if (cand < dn) dn = cand;
            }
            if (dn < d0) {
                worst = std::max(worst, (d0 - dn) / std::max(d0, kMinDist));
                d[c] = dn;
            }
        };
        for (std::size_t c = 0; c < z(m.n_cells); ++c) relax(c);
        for (std::size_t c = z(m.n_cells); c-- > 0;) relax(c);
        if (worst <= rel_tol) break;
    }
    return d;
}

}  // namespace cfdg
