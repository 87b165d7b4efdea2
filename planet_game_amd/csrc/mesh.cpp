// Host construction of the icosahedral SCVT Voronoi mesh, bit-identical to
// the reference's make_icosphere (sim/planet/mesh/icosphere.cpp:200-419):
// the same floating-point operations in the same order, glibc libm for
// atan2/acos/sqrt/hypot, built with -ffp-contract=off. The mesh is set-up
// data: it is built once per level on the host and uploaded.
#include "internal.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>
#include <unordered_map>
#include <utility>

namespace psim {
namespace {

// Vector arithmetic of sim/core/math/vec3d.hpp:8-52.
struct P3 {
    double x = 0.0, y = 0.0, z = 0.0;
};
P3 operator+(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
P3 operator-(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
P3 operator*(P3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
P3 cross(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
P3 normalized(P3 a) {
    const double len = std::sqrt(dot(a, a));
    if (!(len > 0.0) || !std::isfinite(len)) throw std::invalid_argument("cannot normalize a zero or non-finite vector");
    return {a.x / len, a.y / len, a.z / len};
}

using Face = std::array<uint32_t, 3>;

// Spherical triangle area on the unit sphere (icosphere.cpp:53-58).
double spherical_area(P3 a, P3 b, P3 c) {
    return 2.0 * std::atan2(std::abs(dot(a, cross(b, c))), 1.0 + dot(a, b) + dot(b, c) + dot(c, a));
}

// Circumcentre corner of a primal triangle (icosphere.cpp:77-84).
P3 corner_of(P3 a, P3 b, P3 c) {
    P3 n = normalized(cross(b - a, c - a));
    return dot(n, a + b + c) < 0.0 ? n * -1.0 : n;
}

uint64_t undirected(uint32_t a, uint32_t b) {
    return a < b ? (uint64_t{a} << 32U) | b : (uint64_t{b} << 32U) | a;
}

// One Lloyd sweep set per subdivision level (icosphere.cpp:93-118): each
// generator moves to the area-weighted centroid of its kites.
void centroidal_sweeps(std::vector<P3>& v, const std::vector<Face>& faces) {
    std::vector<P3> sum(v.size());
    for (int sweep = 0; sweep < 20; ++sweep) {
        std::fill(sum.begin(), sum.end(), P3{});
        for (const Face& f : faces) {
            const P3 cc = corner_of(v[f[0]], v[f[1]], v[f[2]]);
            for (uint32_t i = 0; i < 3; ++i) {
                const uint32_t g = f[i];
                for (uint32_t other : {f[(i + 1) % 3], f[(i + 2) % 3]}) {
                    const P3 mid = normalized(v[g] + v[other]);
                    const double kite = spherical_area(v[g], mid, cc);
                    sum[g] = sum[g] + normalized(v[g] + mid + cc) * kite;
                }
            }
        }
        for (size_t i = 0; i < v.size(); ++i) v[i] = normalized(sum[i]);
    }
}

}  // namespace

HostMesh build_icosphere(uint32_t level, double radius) {
    if (!(radius > 0.0) || !std::isfinite(radius)) throw std::invalid_argument("icosphere radius must be finite and positive");
    if (level > 12) throw std::overflow_error("icosphere IDs exceed their 32-bit representation");
    const double g = 1.6180339887498948482;
    std::vector<P3> v{{-1.0, g, 0.0}, {1.0, g, 0.0}, {-1.0, -g, 0.0}, {1.0, -g, 0.0},
                      {0.0, -1.0, g}, {0.0, 1.0, g}, {0.0, -1.0, -g}, {0.0, 1.0, -g},
                      {g, 0.0, -1.0}, {g, 0.0, 1.0}, {-g, 0.0, -1.0}, {-g, 0.0, 1.0}};
    for (P3& p : v) p = normalized(p);
    std::vector<Face> faces{{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                            {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                            {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                            {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
    const size_t cells_expected = 10 * (size_t{1} << (2 * level)) + 2;
    v.reserve(cells_expected);
    for (uint32_t l = 0; l < level; ++l) {
        std::unordered_map<uint64_t, uint32_t> split;
        split.reserve(faces.size() * 2);
        auto mid = [&](uint32_t a, uint32_t b) {
            auto [it, inserted] = split.emplace(undirected(a, b), static_cast<uint32_t>(v.size()));
            if (inserted) v.push_back(normalized(v[a] + v[b]));
            return it->second;
        };
        std::vector<Face> finer;
        finer.reserve(faces.size() * 4);
        for (const Face& f : faces) {
            const uint32_t a = mid(f[0], f[1]), b = mid(f[1], f[2]), c = mid(f[2], f[0]);
            finer.push_back({f[0], a, c});
            finer.push_back({f[1], b, a});
            finer.push_back({f[2], c, b});
            finer.push_back({a, b, c});
        }
        faces = std::move(finer);
        centroidal_sweeps(v, faces);
    }
    const uint32_t n = static_cast<uint32_t>(v.size());
    if (n != cells_expected) throw std::logic_error("icosphere subdivision produced unexpected counts");

    // Recursive face order numbers the cells (icosphere.cpp:265-279).
    const uint32_t unset = 0xFFFFFFFFU;
    std::vector<uint32_t> cell_id(n, unset), vertex_of_cell;
    vertex_of_cell.reserve(n);
    for (const Face& f : faces)
        for (uint32_t s : f)
            if (cell_id[s] == unset) {
                cell_id[s] = static_cast<uint32_t>(vertex_of_cell.size());
                vertex_of_cell.push_back(s);
            }
    // Corner k is face k's circumcentre; every vertex lists its faces.
    std::vector<P3> corner(faces.size());
    std::vector<uint32_t> vf_start(n + 1, 0), vf(faces.size() * 3);
    for (const Face& f : faces)
        for (uint32_t s : f) ++vf_start[s + 1];
    for (uint32_t i = 0; i < n; ++i) vf_start[i + 1] += vf_start[i];
    {
        std::vector<uint32_t> fill(vf_start.begin(), vf_start.end() - 1);
        for (uint32_t k = 0; k < faces.size(); ++k) {
            const Face& f = faces[k];
            corner[k] = corner_of(v[f[0]], v[f[1]], v[f[2]]);
            for (uint32_t s : f) vf[fill[s]++] = k;
        }
    }

    HostMesh m;
    m.level = level;
    m.radius = radius;
    m.cells = static_cast<int32_t>(n);
    m.center.resize(3 * size_t{n});
    m.east.resize(3 * size_t{n});
    m.north.resize(3 * size_t{n});
    std::vector<P3> up(n), east(n), north(n);
    // Local tangent basis (local_tangent_basis.cpp:8-24).
    for (uint32_t c = 0; c < n; ++c) {
        up[c] = normalized(v[vertex_of_cell[c]]);
        const double h = std::hypot(up[c].x, up[c].y);
        east[c] = h > 1.0e-12 ? P3{-up[c].y / h, up[c].x / h, 0.0} : P3{0.0, 1.0, 0.0};
        north[c] = normalized(cross(up[c], east[c]));
        const P3* vs[3] = {&up[c], &east[c], &north[c]};
        std::vector<double>* out[3] = {&m.center, &m.east, &m.north};
        for (int q = 0; q < 3; ++q) {
            (*out[q])[3 * c] = vs[q]->x;
            (*out[q])[3 * c + 1] = vs[q]->y;
            (*out[q])[3 * c + 2] = vs[q]->z;
        }
    }

    // Cell polygons, areas, edges and normals (icosphere.cpp:315-397).
    m.area.resize(n);
    m.edge_offset.resize(n);
    m.edge_count.resize(n);
    const size_t directed = 6 * size_t{n} - 12;
    m.neighbor.reserve(directed);
    m.edge.reserve(directed);
    m.normal_east.reserve(directed);
    m.normal_north.reserve(directed);
    std::unordered_map<uint64_t, uint32_t> edge_index;   // undirected pair -> edge id
    edge_index.reserve(directed);
    std::vector<uint8_t> uses;
    for (uint32_t c = 0; c < n; ++c) {
        const uint32_t src = vertex_of_cell[c];
        const uint32_t deg = vf_start[src + 1] - vf_start[src];
        if (deg != 5 && deg != 6) throw std::logic_error("dual cell is neither a pentagon nor a hexagon");
        std::array<std::pair<double, uint32_t>, 6> ring;
        for (uint32_t q = 0; q < deg; ++q) {
            const uint32_t k = vf[vf_start[src] + q];
            const P3 t = corner[k] - up[c] * dot(corner[k], up[c]);
            ring[q] = {std::atan2(dot(t, north[c]), dot(t, east[c])), k};
        }
        std::sort(ring.begin(), ring.begin() + deg);   // by angle, then by corner index
        m.edge_offset[c] = static_cast<uint32_t>(m.neighbor.size());
        m.edge_count[c] = static_cast<uint8_t>(deg);
        double acc = 0.0, lost = 0.0;   // Kahan-compensated area
        for (uint32_t q = 0; q < deg; ++q) {
            const uint32_t k0 = ring[q].second, k1 = ring[(q + 1) % deg].second;
            const double piece = spherical_area(up[c], corner[k0], corner[k1]) - lost;
            const double next = acc + piece;
            lost = (next - acc) - piece;
            acc = next;
            // The neighbour is the other vertex shared by the two faces.
            uint32_t shared = unset;
            for (uint32_t s : faces[k0]) {
                if (s != src && (s == faces[k1][0] || s == faces[k1][1] || s == faces[k1][2])) {
                    shared = s;
                    break;
                }
            }
            if (shared == unset) throw std::logic_error("consecutive dual corners do not share a primal edge");
            const uint32_t nb = cell_id[shared];
            const double len = std::acos(std::clamp(dot(corner[k0], corner[k1]), -1.0, 1.0)) * radius;
            const double dist = std::acos(std::clamp(dot(up[c], up[nb]), -1.0, 1.0)) * radius;
            auto [it, inserted] = edge_index.emplace(undirected(c, nb), static_cast<uint32_t>(m.first.size()));
            const uint32_t e = it->second;
            if (inserted) {
                m.first.push_back(std::min(c, nb));
                m.second.push_back(std::max(c, nb));
                m.length.push_back(len);
                m.distance.push_back(dist);
                uses.push_back(1);
            } else {
                if (uses[e] != 1) throw std::logic_error("dual mesh contains a non-manifold edge");
                ++uses[e];
                if (m.length[e] != len || m.distance[e] != dist) throw std::logic_error("dual edge geometry is not reciprocal");
            }
            P3 conormal = normalized(cross(corner[k0], corner[k1]));
            if (dot(conormal, up[c]) > 0.0) conormal = conormal * -1.0;
            const P3 out = normalized(conormal - up[c] * dot(conormal, up[c]));
            m.neighbor.push_back(nb);
            m.edge.push_back(e);
            m.normal_east.push_back(static_cast<float>(dot(out, east[c])));
            m.normal_north.push_back(static_cast<float>(dot(out, north[c])));
        }
        m.area[c] = acc * radius * radius;
    }
    for (uint8_t u : uses)
        if (u != 2) throw std::logic_error("dual mesh contains an open boundary edge");
    m.edges = static_cast<int32_t>(m.first.size());
    m.directed = static_cast<int32_t>(m.neighbor.size());

    // Transport graph: w = l / d per entry, row sums in entry order.
    m.row.resize(size_t{n} + 1);
    m.weight.resize(m.neighbor.size());
    m.weight_sum.resize(n);
    m.row[0] = 0;
    for (uint32_t c = 0; c < n; ++c) {
        double s = 0.0;
        const uint32_t begin = m.edge_offset[c], end = begin + m.edge_count[c];
        for (uint32_t k = begin; k < end; ++k) {
            m.weight[k] = m.length[m.edge[k]] / m.distance[m.edge[k]];
            s += m.weight[k];
        }
        m.weight_sum[c] = s;
        m.row[c + 1] = static_cast<int32_t>(end);
    }
    return m;
}

HostGraph graph_of(const HostMesh& m) {
    HostGraph g;
    g.n = m.cells;
    g.area = m.area;
    g.row = m.row;
    g.col.assign(m.neighbor.begin(), m.neighbor.end());
    g.weight = m.weight;
    g.weight_sum = m.weight_sum;
    return g;
}

// Each fine cell belongs to the coarse cell whose centre lies nearest on the
// sphere (largest dot product of unit centres). The search hill-climbs the
// coarse mesh from the group of the previous fine cell: fine cells are
// numbered along a space-filling order, so the climb is a step or two. One
// climbing step scans the current cell, then its neighbours in cell-edge
// order, and moves to the first strictly better one than all before it; the
// climb ends at a cell no neighbour improves on. Where two centres tie, the
// start cell and that scan order decide, so both follow the reference
// exactly (heat_transport.cpp:485-503); the dot product sums x, y, z in that
// order (vec3d.hpp:30-32).
HostAgglomeration agglomerate(const HostMesh& fine) {
    HostAgglomeration a;
    const size_t n = static_cast<size_t>(fine.cells);
    a.group.resize(n);
    if (fine.level == 0) {
        a.coarse = fine;
        a.area = fine.area;
        for (size_t c = 0; c < n; ++c) a.group[c] = static_cast<uint32_t>(c);
        return a;
    }
    a.coarse = build_icosphere(fine.level - 1, fine.radius);
    const HostMesh& cm = a.coarse;
    a.area.assign(static_cast<size_t>(cm.cells), 0.0);
    const auto closeness = [&](const double* p, uint32_t q) {
        const double* c = &cm.center[3 * static_cast<size_t>(q)];
        return p[0] * c[0] + p[1] * c[1] + p[2] * c[2];
    };
    uint32_t at = 0;
    for (size_t c = 0; c < n; ++c) {
        const double* p = &fine.center[3 * c];
        for (;;) {
            uint32_t best = at;
            double best_value = closeness(p, at);
            const uint32_t begin = cm.edge_offset[at], end = begin + cm.edge_count[at];
            for (uint32_t k = begin; k < end; ++k) {
                const uint32_t q = cm.neighbor[k];
                const double value = closeness(p, q);
                if (value > best_value) {
                    best_value = value;
                    best = q;
                }
            }
            if (best == at) break;
            at = best;
        }
        a.group[c] = at;
        a.area[at] += fine.area[c];
    }
    for (const double area : a.area)
        if (!(area > 0.0)) throw std::logic_error("a coarse transport cell received no fine cell");
    return a;
}

}  // namespace psim
