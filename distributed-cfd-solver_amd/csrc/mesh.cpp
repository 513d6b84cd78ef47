// Synthetic mesh generation, geometry, RCM renumbering and the 3-zone face
// order for a single-rank mesh.  See mesh.hpp for the contract.
#include "mesh.hpp"

#include "geometry.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <queue>
#include <sstream>
#include <stdexcept>
#include <utility>
#include <vector>

namespace cfdg {

namespace {

// Face tables of the CGNS volume elements used here: local node lists whose
// right-hand-rule normal points out of a positively oriented element
// (include/cfd/mesh/cgnstables.hpp:45-75); geometry.hpp holds the same
// tables for the cell metrics.
constexpr int kTetFaces[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
constexpr int kHexFaces[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5},
                                 {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}};

inline std::uint64_t splitmix64(std::uint64_t x) noexcept {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// Uniform in [-1, 1), a pure function of (seed, stream) so that the node
// perturbation does not depend on traversal order.
inline double uniform_pm1(std::uint64_t seed, std::uint64_t stream) noexcept {
    const std::uint64_t r = splitmix64(seed * 0xD1B54A32D192ED03ULL ^ splitmix64(stream));
    return static_cast<double>(r >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

struct HalfFace {
    std::array<std::int32_t, 4> key;  // sorted node ids, INT32_MAX padding
    std::int32_t cell;
    std::int8_t lface;
};

}  // namespace

void compute_geometry(Mesh& m) {
    const std::size_t nc = static_cast<std::size_t>(m.n_cells);
    const std::size_t nf = static_cast<std::size_t>(m.n_faces);
    m.cell_cx.resize(nc);
    m.cell_cy.resize(nc);
    m.cell_cz.resize(nc);
    m.cell_vol.resize(nc);
    double bx[8], by[8], bz[8];
    for (std::size_t c = 0; c < nc; ++c) {
        const int b = m.cell_nodes_off[c], e = m.cell_nodes_off[c + 1];
        for (int k = b; k < e; ++k) {
            const auto n = static_cast<std::size_t>(m.cell_nodes[static_cast<std::size_t>(k)]);
            bx[k - b] = m.node_x[n];
            by[k - b] = m.node_y[n];
            bz[k - b] = m.node_z[n];
        }
        const int type = (m.cell_type[c] == static_cast<std::uint8_t>(CellKind::Tet)) ? 0 : 3;
        const geom::CellMetrics cm = geom::cell_metrics(type, bx, by, bz);
        if (!(cm.vol > 1e-15)) {
            std::ostringstream os;
            os << "degenerate cell " << c << " volume " << cm.vol;
            throw std::runtime_error(os.str());
        }
        m.cell_cx[c] = cm.cx;
        m.cell_cy[c] = cm.cy;
        m.cell_cz[c] = cm.cz;
        m.cell_vol[c] = cm.vol;
    }

    m.face_cx.resize(nf);
    m.face_cy.resize(nf);
    m.face_cz.resize(nf);
    m.face_nx.resize(nf);
    m.face_ny.resize(nf);
    m.face_nz.resize(nf);
    m.face_area.resize(nf);
    for (std::size_t f = 0; f < nf; ++f) {
        const int b = m.face_nodes_off[f], e = m.face_nodes_off[f + 1];
        const int nn = e - b;
        if (nn < 3 || nn > 8) {
            std::ostringstream os;
            os << "face " << f << " has " << nn << " nodes";
            throw std::runtime_error(os.str());
        }
        for (int k = 0; k < nn; ++k) {
            const auto n = static_cast<std::size_t>(m.face_nodes[static_cast<std::size_t>(b + k)]);
            bx[k] = m.node_x[n];
            by[k] = m.node_y[n];
            bz[k] = m.node_z[n];
        }
        const geom::FaceMetrics fm = geom::face_metrics(nn, bx, by, bz);
        if (!(fm.area > 1e-15)) {
            std::ostringstream os;
            os << "degenerate face " << f << " area " << fm.area;
            throw std::runtime_error(os.str());
        }
        m.face_cx[f] = fm.cx;
        m.face_cy[f] = fm.cy;
        m.face_cz[f] = fm.cz;
        m.face_nx[f] = fm.nx;
        m.face_ny[f] = fm.ny;
        m.face_nz[f] = fm.nz;
        m.face_area[f] = fm.area;
    }
}

Mesh generate_mesh(const MeshSpec& s) {
    if (s.ni < 1 || s.nj < 1 || s.nk < 1) throw std::runtime_error("mesh: ni, nj, nk must be >= 1");
    if (s.kind != 0 && s.kind != 1) throw std::runtime_error("mesh: kind must be 0 (hex) or 1 (tet)");
    if (!(s.x1 > s.x0 && s.y1 > s.y0 && s.z1 > s.z0)) throw std::runtime_error("mesh: empty box");
    if (!(s.y_ratio > 0.0)) throw std::runtime_error("mesh: y_ratio must be > 0");
    if (s.jitter < 0.0 || s.jitter >= 0.5) throw std::runtime_error("mesh: jitter must be in [0, 0.5)");

    const long long nhex = 1LL * s.ni * s.nj * s.nk;
    const long long ncell_ll = (s.kind == 0) ? nhex : 6 * nhex;
    const long long nnode_ll = 1LL * (s.ni + 1) * (s.nj + 1) * (s.nk + 1);
    if (ncell_ll > std::numeric_limits<std::int32_t>::max() / 8 || nnode_ll > std::numeric_limits<std::int32_t>::max())
        throw std::runtime_error("mesh: too many cells for 32-bit local indices");

    Mesh m;
    const int NI = s.ni + 1, NJ = s.nj + 1, NK = s.nk + 1;
    m.n_nodes = static_cast<std::int32_t>(nnode_ll);
    m.n_cells = static_cast<std::int32_t>(ncell_ll);

    // --- node coordinates ----------------------------------------------------
    std::vector<double> xs(static_cast<std::size_t>(NI)), ys(static_cast<std::size_t>(NJ)), zs(static_cast<std::size_t>(NK));
    for (int i = 0; i < NI; ++i) xs[static_cast<std::size_t>(i)] = s.x0 + (s.x1 - s.x0) * (static_cast<double>(i) / s.ni);
    for (int k = 0; k < NK; ++k) zs[static_cast<std::size_t>(k)] = s.z0 + (s.z1 - s.z0) * (static_cast<double>(k) / s.nk);
    for (int j = 0; j < NJ; ++j) {
        double t = static_cast<double>(j) / s.nj;
        if (s.y_ratio != 1.0) t = (std::pow(s.y_ratio, j) - 1.0) / (std::pow(s.y_ratio, s.nj) - 1.0);
        ys[static_cast<std::size_t>(j)] = s.y0 + (s.y1 - s.y0) * t;
    }
    ys[static_cast<std::size_t>(s.nj)] = s.y1;

    const auto spacing = [](const std::vector<double>& v, int i) {
        const int n = static_cast<int>(v.size()) - 1;
        double h = std::numeric_limits<double>::max();
        if (i > 0) h = std::min(h, v[static_cast<std::size_t>(i)] - v[static_cast<std::size_t>(i - 1)]);
        if (i < n) h = std::min(h, v[static_cast<std::size_t>(i + 1)] - v[static_cast<std::size_t>(i)]);
        return h;
    };

    m.node_x.resize(static_cast<std::size_t>(m.n_nodes));
    m.node_y.resize(static_cast<std::size_t>(m.n_nodes));
    m.node_z.resize(static_cast<std::size_t>(m.n_nodes));
    m.node_gid.resize(static_cast<std::size_t>(m.n_nodes));
    const double tan_ramp = std::tan(s.ramp_deg * (3.14159265358979323846 / 180.0));
    for (int k = 0; k < NK; ++k) {
        for (int j = 0; j < NJ; ++j) {
            for (int i = 0; i < NI; ++i) {
                const std::size_t n = static_cast<std::size_t>(i + NI * (j + NJ * k));
                double x = xs[static_cast<std::size_t>(i)], y = ys[static_cast<std::size_t>(j)], z = zs[static_cast<std::size_t>(k)];
                const bool interior = i > 0 && i < s.ni && j > 0 && j < s.nj && k > 0 && k < s.nk;
                if (interior && s.jitter > 0.0) {
                    x += s.jitter * spacing(xs, i) * uniform_pm1(s.seed, 3 * n + 0);
                    y += s.jitter * spacing(ys, j) * uniform_pm1(s.seed, 3 * n + 1);
                    z += s.jitter * spacing(zs, k) * uniform_pm1(s.seed, 3 * n + 2);
                }
                const double eta = (ys[static_cast<std::size_t>(j)] - s.y0) / (s.y1 - s.y0);
                if (s.ramp_deg != 0.0 && x > s.ramp_x) {
                    y += (x - s.ramp_x) * tan_ramp * (1.0 - static_cast<double>(j) / s.nj);
                }
                if (s.bump_h != 0.0) {
                    const double xi = (x - s.bump_x) / s.bump_w;
                    y += s.bump_h * std::exp(-xi * xi) * (1.0 - eta);
                }
                m.node_x[n] = x;
                m.node_y[n] = y;
                m.node_z[n] = z;
                m.node_gid[n] = static_cast<std::int64_t>(n);
            }
        }
    }

    // --- cells -----------------------------------------------------------------
    const auto nid = [&](int i, int j, int k) { return static_cast<std::int32_t>(i + NI * (j + NJ * k)); };
    const int npc = (s.kind == 0) ? 8 : 4;
    m.cell_type.assign(static_cast<std::size_t>(m.n_cells), static_cast<std::uint8_t>(s.kind == 0 ? CellKind::Hexa : CellKind::Tet));
    m.cell_gid.resize(static_cast<std::size_t>(m.n_cells));
    m.cell_nodes_off.resize(static_cast<std::size_t>(m.n_cells) + 1);
    m.cell_nodes.resize(static_cast<std::size_t>(m.n_cells) * static_cast<std::size_t>(npc));
    for (std::int32_t c = 0; c <= m.n_cells; ++c) m.cell_nodes_off[static_cast<std::size_t>(c)] = c * npc;
    {
        static constexpr int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        std::size_t c = 0;
        for (int k = 0; k < s.nk; ++k) {
            for (int j = 0; j < s.nj; ++j) {
                for (int i = 0; i < s.ni; ++i) {
                    if (s.kind == 0) {
                        const std::int32_t h[8] = {nid(i, j, k),         nid(i + 1, j, k),     nid(i + 1, j + 1, k),
                                                   nid(i, j + 1, k),     nid(i, j, k + 1),     nid(i + 1, j, k + 1),
                                                   nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)};
                        std::copy(h, h + 8, m.cell_nodes.begin() + static_cast<std::ptrdiff_t>(8 * c));
                        m.cell_gid[c] = static_cast<std::int64_t>(c);
                        ++c;
                    } else {
                        // Kuhn split: the six monotone lattice paths from corner (0,0,0) to (1,1,1).
                        for (int t = 0; t < 6; ++t) {
                            int o[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 1}};
                            o[1][kPerm[t][0]] = 1;
                            o[2][kPerm[t][0]] = 1;
                            o[2][kPerm[t][1]] = 1;
                            std::int32_t tn[4];
                            for (int v = 0; v < 4; ++v) tn[v] = nid(i + o[v][0], j + o[v][1], k + o[v][2]);
                            // orientation from the lattice offsets: det(o1-o0, o2-o0, o3-o0)
                            const int det = o[1][0] * (o[2][1] * o[3][2] - o[2][2] * o[3][1]) -
                                            o[1][1] * (o[2][0] * o[3][2] - o[2][2] * o[3][0]) +
                                            o[1][2] * (o[2][0] * o[3][1] - o[2][1] * o[3][0]);
                            if (det < 0) std::swap(tn[1], tn[2]);
                            std::copy(tn, tn + 4, m.cell_nodes.begin() + static_cast<std::ptrdiff_t>(4 * c));
                            m.cell_gid[c] = static_cast<std::int64_t>(c);
                            ++c;
                        }
                    }
                }
            }
        }
    }

    // --- unique faces ------------------------------------------------------------
    const int fpc = (s.kind == 0) ? 6 : 4;
    std::vector<HalfFace> hf(static_cast<std::size_t>(m.n_cells) * static_cast<std::size_t>(fpc));
    for (std::int32_t c = 0; c < m.n_cells; ++c) {
        const std::int32_t* cn = m.cell_nodes.data() + static_cast<std::size_t>(c) * static_cast<std::size_t>(npc);
        for (int f = 0; f < fpc; ++f) {
            HalfFace& h = hf[static_cast<std::size_t>(c) * static_cast<std::size_t>(fpc) + static_cast<std::size_t>(f)];
            h.key = {std::numeric_limits<std::int32_t>::max(), std::numeric_limits<std::int32_t>::max(),
                     std::numeric_limits<std::int32_t>::max(), std::numeric_limits<std::int32_t>::max()};
            const int nn = (s.kind == 0) ? 4 : 3;
            for (int v = 0; v < nn; ++v) h.key[static_cast<std::size_t>(v)] = (s.kind == 0) ? cn[kHexFaces[f][v]] : cn[kTetFaces[f][v]];
            std::sort(h.key.begin(), h.key.begin() + nn);
            h.cell = c;
            h.lface = static_cast<std::int8_t>(f);
        }
    }
    std::sort(hf.begin(), hf.end(), [](const HalfFace& a, const HalfFace& b) {
        if (a.key != b.key) return a.key < b.key;
        return a.cell < b.cell;
    });

    struct FaceRec {
        std::int32_t owner, neigh, patch;
        std::int8_t lface;
    };
    std::vector<FaceRec> faces;
    faces.reserve(hf.size() / 2 + static_cast<std::size_t>(2 * (s.ni * s.nj + s.nj * s.nk + s.ni * s.nk)) * 2);

    const auto node_ijk = [&](std::int32_t n, int& i, int& j, int& k) {
        i = n % NI;
        j = (n / NI) % NJ;
        k = n / (NI * NJ);
    };
    std::vector<std::string> names = {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"};
    if (s.split_ymin) names.push_back("ymin_lo");
    const auto patch_of = [&](const HalfFace& h) -> std::int32_t {
        const int nn = (s.kind == 0) ? 4 : 3;
        bool all[6] = {true, true, true, true, true, true};
        double xc = 0.0;
        for (int v = 0; v < nn; ++v) {
            int i, j, k;
            node_ijk(h.key[static_cast<std::size_t>(v)], i, j, k);
            all[0] = all[0] && i == 0;
            all[1] = all[1] && i == s.ni;
            all[2] = all[2] && j == 0;
            all[3] = all[3] && j == s.nj;
            all[4] = all[4] && k == 0;
            all[5] = all[5] && k == s.nk;
            xc += m.node_x[static_cast<std::size_t>(h.key[static_cast<std::size_t>(v)])];
        }
        xc /= nn;
        for (int p = 0; p < 6; ++p) {
            if (all[p]) {
                if (p == 2 && s.split_ymin && xc < s.split_ymin_x) return 6;
                return p;
            }
        }
        throw std::runtime_error("mesh: boundary face not on the box surface");
    };

    for (std::size_t a = 0; a < hf.size();) {
        if (a + 1 < hf.size() && hf[a + 1].key == hf[a].key) {
            faces.push_back({hf[a].cell, hf[a + 1].cell, -1, hf[a].lface});  // hf sorted by cell within a key
            a += 2;
        } else {
            faces.push_back({hf[a].cell, -1, patch_of(hf[a]), hf[a].lface});
            a += 1;
        }
    }
    std::vector<HalfFace>().swap(hf);

    std::sort(faces.begin(), faces.end(), [](const FaceRec& a, const FaceRec& b) {
        const int za = a.neigh < 0 ? 2 : 0, zb = b.neigh < 0 ? 2 : 0;
        if (za != zb) return za < zb;
        if (za == 2) {
            if (a.patch != b.patch) return a.patch < b.patch;
            return a.owner < b.owner;
        }
        if (a.owner != b.owner) return a.owner < b.owner;
        return a.neigh < b.neigh;
    });

    m.n_faces = static_cast<std::int32_t>(faces.size());
    const std::size_t nf = faces.size();
    m.face_owner.resize(nf);
    m.face_neigh.resize(nf);
    m.face_patch.resize(nf);
    m.face_type.resize(nf);
    m.face_nodes_off.resize(nf + 1);
    const int fnn = (s.kind == 0) ? 4 : 3;
    m.face_nodes.resize(nf * static_cast<std::size_t>(fnn));
    m.n_inner_faces = 0;
    for (std::size_t f = 0; f < nf; ++f) {
        const FaceRec& r = faces[f];
        m.face_owner[f] = r.owner;
        m.face_neigh[f] = r.neigh;
        m.face_patch[f] = r.patch;
        m.face_type[f] = static_cast<std::uint8_t>(s.kind == 0 ? CellKind::Quad : CellKind::Tri);
        m.face_nodes_off[f] = static_cast<std::int32_t>(f) * fnn;
        const std::int32_t* cn = m.cell_nodes.data() + static_cast<std::size_t>(r.owner) * static_cast<std::size_t>(npc);
        for (int v = 0; v < fnn; ++v) {
            m.face_nodes[f * static_cast<std::size_t>(fnn) + static_cast<std::size_t>(v)] =
                (s.kind == 0) ? cn[kHexFaces[r.lface][v]] : cn[kTetFaces[r.lface][v]];
        }
        if (r.neigh >= 0) ++m.n_inner_faces;
    }
    m.face_nodes_off[nf] = static_cast<std::int32_t>(nf) * fnn;
    m.n_internal_faces = m.n_inner_faces;
    m.patch_names = names;
    m.patch_cgns_types.assign(names.size(), "UserDefined");

    compute_geometry(m);

    m.bbox_lo[0] = m.bbox_lo[1] = m.bbox_lo[2] = std::numeric_limits<double>::max();
    m.bbox_hi[0] = m.bbox_hi[1] = m.bbox_hi[2] = -std::numeric_limits<double>::max();
    for (std::size_t n = 0; n < m.node_x.size(); ++n) {
        m.bbox_lo[0] = std::min(m.bbox_lo[0], m.node_x[n]);
        m.bbox_lo[1] = std::min(m.bbox_lo[1], m.node_y[n]);
        m.bbox_lo[2] = std::min(m.bbox_lo[2], m.node_z[n]);
        m.bbox_hi[0] = std::max(m.bbox_hi[0], m.node_x[n]);
        m.bbox_hi[1] = std::max(m.bbox_hi[1], m.node_y[n]);
        m.bbox_hi[2] = std::max(m.bbox_hi[2], m.node_z[n]);
    }

    if (s.reorder == 1) {
        apply_cell_permutation(m, rcm_order(m));
    }
    const std::string err = check_mesh(m);
    if (!err.empty()) throw std::runtime_error("mesh: " + err);
    return m;
}

std::vector<std::int32_t> rcm_order(const Mesh& m) {
    const std::int32_t n = m.n_cells;
    const std::size_t nsz = static_cast<std::size_t>(n);
    // Dual graph over owned cells, adjacency appended in face order.
    std::vector<std::int32_t> deg(nsz, 0), off(nsz + 1, 0), adj;
    for (std::int32_t f = 0; f < m.n_faces; ++f) {
        const std::int32_t u = m.face_owner[static_cast<std::size_t>(f)], v = m.face_neigh[static_cast<std::size_t>(f)];
        if (v >= 0 && v < n) {
            ++deg[static_cast<std::size_t>(u)];
            ++deg[static_cast<std::size_t>(v)];
        }
    }
    for (std::size_t i = 0; i < nsz; ++i) off[i + 1] = off[i] + deg[i];
    adj.resize(static_cast<std::size_t>(off[nsz]));
    std::vector<std::int32_t> cur(off.begin(), off.end() - 1);
    for (std::int32_t f = 0; f < m.n_faces; ++f) {
        const std::int32_t u = m.face_owner[static_cast<std::size_t>(f)], v = m.face_neigh[static_cast<std::size_t>(f)];
        if (v >= 0 && v < n) {
            adj[static_cast<std::size_t>(cur[static_cast<std::size_t>(u)]++)] = v;
            adj[static_cast<std::size_t>(cur[static_cast<std::size_t>(v)]++)] = u;
        }
    }

    std::vector<std::uint8_t> visited(nsz, 0);
    // Pseudo-peripheral start: up to three BFS passes, each restarting from the
    // deepest node of the previous one (lowest degree among the deepest).
    const auto peripheral = [&](std::int32_t root) {
        std::int32_t best = root;
        int max_depth = 0;
        std::vector<std::uint8_t> seen(nsz);
        std::vector<std::pair<std::int32_t, int>> q;
        for (int pass = 0; pass < 3; ++pass) {
            std::fill(seen.begin(), seen.end(), 0);
            q.clear();
            q.emplace_back(best, 0);
            seen[static_cast<std::size_t>(best)] = 1;
            std::int32_t far = best;
            int deepest = 0;
            std::int32_t min_deg = deg[static_cast<std::size_t>(best)];
            for (std::size_t h = 0; h < q.size(); ++h) {
                const auto [u, d] = q[h];
                if (d > deepest || (d == deepest && deg[static_cast<std::size_t>(u)] < min_deg)) {
                    deepest = d;
                    far = u;
                    min_deg = deg[static_cast<std::size_t>(u)];
                }
                for (std::int32_t e = off[static_cast<std::size_t>(u)]; e < off[static_cast<std::size_t>(u) + 1]; ++e) {
                    const std::int32_t v = adj[static_cast<std::size_t>(e)];
                    if (!seen[static_cast<std::size_t>(v)] && !visited[static_cast<std::size_t>(v)]) {
                        seen[static_cast<std::size_t>(v)] = 1;
                        q.emplace_back(v, d + 1);
                    }
                }
            }
            if (deepest <= max_depth) break;
            max_depth = deepest;
            best = far;
        }
        return best;
    };

    std::vector<std::int32_t> order;
    order.reserve(nsz);
    std::vector<std::pair<std::int32_t, std::int32_t>> nbs;
    for (std::int32_t i = 0; i < n; ++i) {
        if (visited[static_cast<std::size_t>(i)]) continue;
        const std::int32_t start = peripheral(i);
        std::size_t head = order.size();
        order.push_back(start);
        visited[static_cast<std::size_t>(start)] = 1;
        while (head < order.size()) {
            const std::int32_t u = order[head++];
            nbs.clear();
            for (std::int32_t e = off[static_cast<std::size_t>(u)]; e < off[static_cast<std::size_t>(u) + 1]; ++e) {
                const std::int32_t v = adj[static_cast<std::size_t>(e)];
                if (!visited[static_cast<std::size_t>(v)]) {
                    visited[static_cast<std::size_t>(v)] = 1;
                    nbs.emplace_back(deg[static_cast<std::size_t>(v)], v);
                }
            }
            std::sort(nbs.begin(), nbs.end());
            for (const auto& pv : nbs) order.push_back(pv.second);
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

void apply_cell_permutation(Mesh& m, const std::vector<std::int32_t>& new2old) {
    const std::size_t nc = static_cast<std::size_t>(m.n_cells);
    if (new2old.size() != nc) throw std::runtime_error("mesh: permutation size mismatch");
    std::vector<std::int32_t> old2new(nc, -1);
    for (std::size_t c = 0; c < nc; ++c) {
        const std::int32_t o = new2old[c];
        if (o < 0 || static_cast<std::size_t>(o) >= nc || old2new[static_cast<std::size_t>(o)] != -1)
            throw std::runtime_error("mesh: invalid permutation");
        old2new[static_cast<std::size_t>(o)] = static_cast<std::int32_t>(c);
    }

    const auto permute = [&](auto& v) {
        auto out = v;
        for (std::size_t c = 0; c < nc; ++c) out[c] = v[static_cast<std::size_t>(new2old[c])];
        v.swap(out);
    };
    permute(m.cell_type);
    permute(m.cell_gid);
    permute(m.cell_cx);
    permute(m.cell_cy);
    permute(m.cell_cz);
    permute(m.cell_vol);
    {
        std::vector<std::int32_t> off(nc + 1, 0), nodes;
        nodes.reserve(m.cell_nodes.size());
        for (std::size_t c = 0; c < nc; ++c) {
            const auto o = static_cast<std::size_t>(new2old[c]);
            for (int k = m.cell_nodes_off[o]; k < m.cell_nodes_off[o + 1]; ++k) nodes.push_back(m.cell_nodes[static_cast<std::size_t>(k)]);
            off[c + 1] = static_cast<std::int32_t>(nodes.size());
        }
        m.cell_nodes_off.swap(off);
        m.cell_nodes.swap(nodes);
    }

    const std::size_t nf = static_cast<std::size_t>(m.n_faces);
    for (std::size_t f = 0; f < nf; ++f) {
        m.face_owner[f] = old2new[static_cast<std::size_t>(m.face_owner[f])];
        if (m.face_neigh[f] >= 0) m.face_neigh[f] = old2new[static_cast<std::size_t>(m.face_neigh[f])];
    }
    std::vector<std::int32_t> perm(nf);
    for (std::size_t f = 0; f < nf; ++f) perm[f] = static_cast<std::int32_t>(f);
    std::sort(perm.begin(), perm.end(), [&](std::int32_t a, std::int32_t b) {
        const auto as = static_cast<std::size_t>(a), bs = static_cast<std::size_t>(b);
        const int za = m.face_neigh[as] < 0 ? 2 : 0, zb = m.face_neigh[bs] < 0 ? 2 : 0;
        if (za != zb) return za < zb;
        if (za == 2) {
            if (m.face_patch[as] != m.face_patch[bs]) return m.face_patch[as] < m.face_patch[bs];
            return m.face_owner[as] < m.face_owner[bs];
        }
        if (m.face_owner[as] != m.face_owner[bs]) return m.face_owner[as] < m.face_owner[bs];
        return m.face_neigh[as] < m.face_neigh[bs];
    });
    const auto fpermute = [&](auto& v) {
        auto out = v;
        for (std::size_t f = 0; f < nf; ++f) out[f] = v[static_cast<std::size_t>(perm[f])];
        v.swap(out);
    };
    fpermute(m.face_owner);
    fpermute(m.face_neigh);
    fpermute(m.face_patch);
    fpermute(m.face_type);
    fpermute(m.face_cx);
    fpermute(m.face_cy);
    fpermute(m.face_cz);
    fpermute(m.face_nx);
    fpermute(m.face_ny);
    fpermute(m.face_nz);
    fpermute(m.face_area);
    {
        std::vector<std::int32_t> off(nf + 1, 0), nodes;
        nodes.reserve(m.face_nodes.size());
        for (std::size_t f = 0; f < nf; ++f) {
            const auto o = static_cast<std::size_t>(perm[f]);
            for (int k = m.face_nodes_off[o]; k < m.face_nodes_off[o + 1]; ++k) nodes.push_back(m.face_nodes[static_cast<std::size_t>(k)]);
            off[f + 1] = static_cast<std::int32_t>(nodes.size());
        }
        m.face_nodes_off.swap(off);
        m.face_nodes.swap(nodes);
    }
    std::int32_t n_inner = 0;
    while (n_inner < m.n_faces && m.face_neigh[static_cast<std::size_t>(n_inner)] >= 0) ++n_inner;
    m.n_inner_faces = n_inner;
    m.n_internal_faces = n_inner;
}

std::string check_mesh(const Mesh& m) {
    std::ostringstream os;
    const std::size_t nc = static_cast<std::size_t>(m.n_cells), nf = static_cast<std::size_t>(m.n_faces);
    if (m.cell_nodes_off.size() != nc + 1 || m.cell_vol.size() != nc || m.cell_cx.size() != nc) return "cell array sizes";
    if (m.face_owner.size() != nf || m.face_neigh.size() != nf || m.face_area.size() != nf || m.face_nodes_off.size() != nf + 1)
        return "face array sizes";
    for (std::size_t f = 0; f < nf; ++f) {
        const std::int32_t o = m.face_owner[f], n = m.face_neigh[f];
        if (o < 0 || o >= m.n_cells) return "face owner out of range";
        const bool inner = static_cast<std::int32_t>(f) < m.n_inner_faces;
        if (inner) {
            if (n < 0 || n >= m.n_cells || n == o) return "inner face neighbour out of range";
            if (m.face_patch[f] != -1) return "inner face carries a patch";
        } else {
            if (n != -1) return "boundary face has a neighbour";
            if (m.face_patch[f] < 0 || m.face_patch[f] >= static_cast<std::int32_t>(m.patch_names.size()))
                return "boundary patch out of range";
            if (static_cast<std::int32_t>(f) > m.n_inner_faces && m.face_patch[f] < m.face_patch[f - 1])
                return "boundary patches not contiguous";
        }
        // outward normal from the owner centroid
        const double d = (m.face_cx[f] - m.cell_cx[static_cast<std::size_t>(o)]) * m.face_nx[f] +
                         (m.face_cy[f] - m.cell_cy[static_cast<std::size_t>(o)]) * m.face_ny[f] +
                         (m.face_cz[f] - m.cell_cz[static_cast<std::size_t>(o)]) * m.face_nz[f];
        if (!(d > 0.0)) {
            os << "face " << f << " normal does not point out of its owner";
            return os.str();
        }
    }
    for (std::size_t c = 0; c < nc; ++c) {
        if (!(m.cell_vol[c] > 1e-15)) return "non-positive cell volume";
    }
    return std::string();
}

}  // namespace cfdg
