#include "setup.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

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
    Csr out;
    out.off.assign(z(m.n_cells) + 1, 0);
    std::vector<i32> stamp(z(m.n_cells), -1);
    for (i32 c = 0; c < m.n_cells; ++c) {
        const std::size_t start = out.idx.size();
        for (i32 k = m.cell_nodes_off[z(c)]; k < m.cell_nodes_off[z(c) + 1]; ++k) {
            const i32 n = m.cell_nodes[z(k)];
            for (i32 j = nc.off[z(n)]; j < nc.off[z(n) + 1]; ++j) {
                const i32 cand = nc.idx[z(j)];
                if (cand == c || stamp[z(cand)] == c) continue;
                stamp[z(cand)] = c;
                out.idx.push_back(cand);
            }
        }
        std::sort(out.idx.begin() + static_cast<std::ptrdiff_t>(start), out.idx.end());
        out.off[z(c) + 1] = static_cast<i32>(out.idx.size());
    }
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

Tiling tile_cells(const Mesh& m, const Csr& ccf, const Csr& cf, int max_cells, int max_faces) {
    const i32 n = m.n_cells;
    Tiling t;
    t.old2new.assign(z(n), -1);
    t.new2old.reserve(z(n));
    t.tile_off.push_back(0);
    std::vector<i32> queue, members;
    std::vector<i32> face_mark(z(m.n_faces), -1);
    i32 tile = 0;
    for (i32 seed = 0; seed < n; ++seed) {
        if (t.old2new[z(seed)] >= 0) continue;
        members.clear();
        queue.clear();
        queue.push_back(seed);
        t.old2new[z(seed)] = -2;  // queued
        int n_faces = 0;
        std::size_t head = 0;
        while (head < queue.size() && static_cast<int>(members.size()) < max_cells) {
            const i32 c = queue[head++];
            // faces this cell would add to the tile
            int add = 0;
            for (i32 e = cf.off[z(c)]; e < cf.off[z(c) + 1]; ++e)
                if (face_mark[z(cf.idx[z(e)])] != tile) ++add;
            if (n_faces + add > max_faces) {
                t.old2new[z(c)] = -1;  // leave for a later tile
                continue;
            }
            n_faces += add;
            for (i32 e = cf.off[z(c)]; e < cf.off[z(c) + 1]; ++e) face_mark[z(cf.idx[z(e)])] = tile;
            members.push_back(c);
            t.old2new[z(c)] = -3;  // member
            for (i32 e = ccf.off[z(c)]; e < ccf.off[z(c) + 1]; ++e) {
                const i32 nb = ccf.idx[z(e)];
                if (t.old2new[z(nb)] == -1) {
                    t.old2new[z(nb)] = -2;
                    queue.push_back(nb);
                }
            }
        }
        for (std::size_t q = head; q < queue.size(); ++q)
            if (t.old2new[z(queue[q])] == -2) t.old2new[z(queue[q])] = -1;
        std::sort(members.begin(), members.end());
        for (const i32 c : members) {
            t.old2new[z(c)] = static_cast<i32>(t.new2old.size());
            t.new2old.push_back(c);
        }
        t.tile_off.push_back(static_cast<i32>(t.new2old.size()));
        ++tile;
    }
    if (static_cast<i32>(t.new2old.size()) != n) throw std::runtime_error("tiling: not every cell assigned");
    return t;
}

TileFaces tile_faces(const Tiling& t, const Csr& dev_cf, const std::vector<i32>& face_owner_dev) {
    TileFaces tf;
    const std::size_t ntiles = t.tile_off.size() - 1;
    tf.tile_face_off.assign(ntiles + 1, 0);
    tf.cell_slot.assign(dev_cf.idx.size(), 0);
    std::vector<i32> local;
    for (std::size_t k = 0; k < ntiles; ++k) {
        local.clear();
        for (i32 c = t.tile_off[k]; c < t.tile_off[k + 1]; ++c)
            for (i32 e = dev_cf.off[z(c)]; e < dev_cf.off[z(c) + 1]; ++e) local.push_back(dev_cf.idx[z(e)]);
        std::sort(local.begin(), local.end());
        local.erase(std::unique(local.begin(), local.end()), local.end());
        for (i32 c = t.tile_off[k]; c < t.tile_off[k + 1]; ++c) {
            for (i32 e = dev_cf.off[z(c)]; e < dev_cf.off[z(c) + 1]; ++e) {
                const i32 f = dev_cf.idx[z(e)];
                const i32 slot = static_cast<i32>(std::lower_bound(local.begin(), local.end(), f) - local.begin());
                tf.cell_slot[z(e)] = (face_owner_dev[z(f)] == c) ? slot : -slot - 1;
            }
        }
        tf.faces.insert(tf.faces.end(), local.begin(), local.end());
        tf.tile_face_off[k + 1] = static_cast<i32>(tf.faces.size());
    }
    return tf;
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

}  // namespace cfdg
