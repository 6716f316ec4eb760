// refcpu core: TEST INFRASTRUCTURE (the oracle). Restates solidfem's CPU path in the
// reference's floating-point order. Built with -ffp-contract=off so no multiply-add is fused
// (the reference refuses mul_add: clippy.toml:8-9). Nothing in the product may include this.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace refcpu {

using u32 = uint32_t;
using u64 = uint64_t;
constexpr u32 FIXED = 0xffffffffu;  // crates/solidfem-elastic/src/assembly.rs:15

// ---------------------------------------------------------------------------------------------
// Mesh: crates/solidfem-mesh/src/{box_mesh,quadratic,boundary,tet}.rs
// ---------------------------------------------------------------------------------------------

struct Vec3 {
    double v[3];
    double &operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
};

// box_mesh.rs:84-90
inline double box_coord(double origin, double size, size_t index, size_t divisions) {
    if (index == divisions) return origin + size;
    return origin + size * ((double)index / (double)divisions);
}

struct LinearMesh {
    std::vector<Vec3> positions;
    std::vector<std::array<u32, 4>> tets;
};

// box_mesh.rs:11-18 -- the six axis orders and whether the walk keeps its orientation
static const int AXIS_ORDERS[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {0, 2, 1}, {2, 1, 0}, {1, 0, 2}};
static const bool AXIS_EVEN[6] = {true, true, true, false, false, false};

// box_mesh.rs:22-80
inline LinearMesh box_tet_mesh(const double origin[3], const double size[3], const int divisions[3]) {
    LinearMesh m;
    size_t dims[3];
    for (int a = 0; a < 3; ++a) {
        if (divisions[a] <= 0) throw std::runtime_error("each axis needs at least one division");
        if (!(size[a] > 0.0)) throw std::runtime_error("each box length must be positive");
        dims[a] = (size_t)divisions[a] + 1;
    }
    auto node_index = [&](size_t i, size_t j, size_t k) { return (u32)(i + dims[0] * (j + dims[1] * k)); };
    m.positions.reserve(dims[0] * dims[1] * dims[2]);
    for (size_t k = 0; k < dims[2]; ++k)
        for (size_t j = 0; j < dims[1]; ++j)
            for (size_t i = 0; i < dims[0]; ++i)
                m.positions.push_back(Vec3{{box_coord(origin[0], size[0], i, divisions[0]),
                                            box_coord(origin[1], size[1], j, divisions[1]),
                                            box_coord(origin[2], size[2], k, divisions[2])}});
    for (size_t k = 0; k < (size_t)divisions[2]; ++k)
        for (size_t j = 0; j < (size_t)divisions[1]; ++j)
            for (size_t i = 0; i < (size_t)divisions[0]; ++i)
                for (int o = 0; o < 6; ++o) {
                    size_t at[3] = {i, j, k};
                    u32 v0 = node_index(at[0], at[1], at[2]);
                    at[AXIS_ORDERS[o][0]] += 1;
                    u32 v1 = node_index(at[0], at[1], at[2]);
                    at[AXIS_ORDERS[o][1]] += 1;
                    u32 v2 = node_index(at[0], at[1], at[2]);
                    at[AXIS_ORDERS[o][2]] += 1;
                    u32 v3 = node_index(at[0], at[1], at[2]);
                    if (AXIS_EVEN[o]) m.tets.push_back({v0, v1, v2, v3});
                    else m.tets.push_back({v0, v2, v1, v3});
                }
    return m;
}

// demo/wasm/src/part.rs:48-82: the grid lines across a tube's wall, `walls` cells through each
// wall and `inner` cells over the hollow's span, measured from 0.
inline std::vector<double> tube_cross_lines(double along, double wall, int walls, int inner) {
    const double hollow = along - 2.0 * wall, step = wall / (double)walls;
    std::vector<double> lines;
    for (int i = 0; i <= walls; ++i) lines.push_back(step * (double)i);
    for (int i = 1; i < inner; ++i) lines.push_back(wall + hollow * (double)i / (double)inner);
    for (int i = 0; i <= walls; ++i) lines.push_back(along - wall + step * (double)i);
    return lines;
}

// demo/wasm/src/part.rs:111-185: a unit-spaced box grid whose tetrahedra are kept unless the
// integer floor of their centroid lies in the hollow, vertices renumbered by first use and
// moved to the grid lines at their rounded unit coordinates. `across` are the two axes of the
// section; `lines[a]` the grid lines of axis a.
inline LinearMesh tube_tet_mesh(const std::vector<double> lines[3], const int across[2], int walls) {
    int cells[3];
    double unit_size[3];
    for (int a = 0; a < 3; ++a) {
        cells[a] = (int)lines[a].size() - 1;
        unit_size[a] = (double)cells[a];
    }
    const double zero[3] = {0.0, 0.0, 0.0};
    const LinearMesh unit = box_tet_mesh(zero, unit_size, cells);
    auto hollow = [&](size_t j, size_t k) {
        return walls > 0 && j >= (size_t)walls && j < (size_t)(cells[across[0]] - walls) && k >= (size_t)walls &&
               k < (size_t)(cells[across[1]] - walls);
    };
    LinearMesh m;
    std::vector<u32> remap(unit.positions.size(), 0xffffffffu);
    for (const auto &tet : unit.tets) {
        double centroid[3] = {0.0, 0.0, 0.0};
        for (u32 n : tet)
            for (int a = 0; a < 3; ++a) centroid[a] = centroid[a] + unit.positions[n][a] * 0.25;
        if (hollow((size_t)std::floor(centroid[across[0]]), (size_t)std::floor(centroid[across[1]]))) continue;
        std::array<u32, 4> out;
        for (int c = 0; c < 4; ++c) {
            u32 &slot = remap[tet[c]];
            if (slot == 0xffffffffu) {
                slot = (u32)m.positions.size();
                const Vec3 &p = unit.positions[tet[c]];
                m.positions.push_back(Vec3{{lines[0][(size_t)std::round(p[0])], lines[1][(size_t)std::round(p[1])],
                                            lines[2][(size_t)std::round(p[2])]}});
            }
            out[c] = slot;
        }
        m.tets.push_back(out);
    }
    return m;
}

// tet.rs:9-16
inline double tet_volume(const Vec3 &a, const Vec3 &b, const Vec3 &c, const Vec3 &d) {
    double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double w[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
    double det = u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) +
                 u[2] * (v[0] * w[1] - v[1] * w[0]);
    return det / 6.0;
}

// tet.rs:59-109 (the volume checks; the other refusals cannot occur for box meshes)
inline void validate_linear(const LinearMesh &m) {
    if (m.tets.empty()) throw std::runtime_error("the mesh has no tetrahedra");
    for (size_t t = 0; t < m.tets.size(); ++t) {
        const auto &tet = m.tets[t];
        double vol = tet_volume(m.positions[tet[0]], m.positions[tet[1]], m.positions[tet[2]], m.positions[tet[3]]);
        if (std::isnan(vol) || vol <= 0.0) throw std::runtime_error("degenerate tetrahedron " + std::to_string(t));
    }
}

// boundary.rs:8-13
static const int FACE_NODES[4][6] = {{1, 2, 3, 5, 9, 8}, {0, 3, 2, 7, 9, 6}, {0, 1, 3, 4, 8, 7}, {0, 2, 1, 6, 5, 4}};
// quadratic.rs:10
static const int EDGE_CORNERS[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

struct BoundaryFace {
    u32 tet;
    uint8_t face;
    u32 nodes[6];
};

struct QuadMesh {
    std::vector<Vec3> nodes;
    std::vector<std::array<u32, 10>> tets;
    std::vector<std::array<u32, 4>> corner_tets;
    size_t vertex_count = 0;
};

// quadratic.rs:24-71
inline QuadMesh quadratic_from_linear(const LinearMesh &m) {
    QuadMesh q;
    q.vertex_count = m.positions.size();
    std::vector<std::pair<u32, u32>> edges;
    edges.reserve(m.tets.size() * 6);
    for (const auto &t : m.tets)
        for (auto &e : EDGE_CORNERS) {
            u32 x = t[e[0]], y = t[e[1]];
            edges.push_back({std::min(x, y), std::max(x, y)});
        }
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    q.nodes = m.positions;
    q.nodes.reserve(q.nodes.size() + edges.size());
    for (auto &e : edges) {
        Vec3 p = q.nodes[e.first], r = q.nodes[e.second];
        q.nodes.push_back(Vec3{{0.5 * (p[0] + r[0]), 0.5 * (p[1] + r[1]), 0.5 * (p[2] + r[2])}});
    }
    q.tets.reserve(m.tets.size());
    for (const auto &t : m.tets) {
        std::array<u32, 10> quad{};
        for (int i = 0; i < 4; ++i) quad[i] = t[i];
        for (int s = 0; s < 6; ++s) {
            u32 x = t[EDGE_CORNERS[s][0]], y = t[EDGE_CORNERS[s][1]];
            auto key = std::make_pair(std::min(x, y), std::max(x, y));
            size_t at = std::lower_bound(edges.begin(), edges.end(), key) - edges.begin();
            quad[4 + s] = (u32)(q.vertex_count + at);
        }
        q.tets.push_back(quad);
    }
    q.corner_tets = m.tets;
    return q;
}

// quadratic.rs:89-124
inline std::vector<BoundaryFace> boundary_faces(const QuadMesh &q) {
    struct Keyed {
        u32 key[3];
        size_t tet;
        int face;
        bool operator<(const Keyed &o) const {
            if (key[0] != o.key[0]) return key[0] < o.key[0];
            if (key[1] != o.key[1]) return key[1] < o.key[1];
            if (key[2] != o.key[2]) return key[2] < o.key[2];
            if (tet != o.tet) return tet < o.tet;
            return face < o.face;
        }
    };
    std::vector<Keyed> keyed;
    keyed.reserve(q.corner_tets.size() * 4);
    for (size_t t = 0; t < q.corner_tets.size(); ++t)
        for (int f = 0; f < 4; ++f) {
            Keyed k;
            for (int s = 0; s < 3; ++s) k.key[s] = q.corner_tets[t][FACE_NODES[f][s]];
            std::sort(k.key, k.key + 3);
            k.tet = t;
            k.face = f;
            keyed.push_back(k);
        }
    std::sort(keyed.begin(), keyed.end());
    std::vector<BoundaryFace> out;
    size_t at = 0;
    while (at < keyed.size()) {
        size_t end = at + 1;
        while (end < keyed.size() && std::equal(keyed[end].key, keyed[end].key + 3, keyed[at].key)) ++end;
        if (end == at + 1) {
            BoundaryFace b;
            b.tet = (u32)keyed[at].tet;
            b.face = (uint8_t)keyed[at].face;
            for (int s = 0; s < 6; ++s) b.nodes[s] = q.tets[keyed[at].tet][FACE_NODES[keyed[at].face][s]];
            out.push_back(b);
        }
        at = end;
    }
    return out;
}

// ---------------------------------------------------------------------------------------------
// Element: crates/solidfem-elastic/src/{shape,quadrature,material,element}.rs
// ---------------------------------------------------------------------------------------------

struct QPoint {
    double natural[3];
    double weight;
};

// quadrature.rs:44-51
static const double A4 = 0.1381966011250105;
static const double B4 = 0.5854101966249685;
static const QPoint FOUR_POINT[4] = {
    {{A4, A4, A4}, 0.25}, {{B4, A4, A4}, 0.25}, {{A4, B4, A4}, 0.25}, {{A4, A4, B4}, 0.25}};
// quadrature.rs:53-77
static const double A14 = 0.3108859192633006;
static const double B14 = 1.0 - 3.0 * A14;
static const double W14A = 0.11268792571801585;
static const double C14 = 0.09273525031089123;
static const double D14 = 1.0 - 3.0 * C14;
static const double W14B = 0.07349304311636195;
static const double E14 = 0.04550370412564965;
static const double F14 = 0.5 - E14;
static const double W14C = 0.042546020777081466;
static const QPoint FOURTEEN_POINT[14] = {
    {{A14, A14, A14}, W14A}, {{B14, A14, A14}, W14A}, {{A14, B14, A14}, W14A}, {{A14, A14, B14}, W14A},
    {{C14, C14, C14}, W14B}, {{D14, C14, C14}, W14B}, {{C14, D14, C14}, W14B}, {{C14, C14, D14}, W14B},
    {{E14, E14, F14}, W14C}, {{E14, F14, E14}, W14C}, {{F14, E14, E14}, W14C}, {{E14, F14, F14}, W14C},
    {{F14, E14, F14}, W14C}, {{F14, F14, E14}, W14C}};
// quadrature.rs:87-94
static const double TA = 0.091576213509771;
static const double TB = 0.445948490915965;
static const double WA = 0.109951743655322;
static const double WB = 0.223381589678011;
static const double TRIANGLE_POINTS[6][3] = {{TA, TA, WA}, {1.0 - 2.0 * TA, TA, WA}, {TA, 1.0 - 2.0 * TA, WA},
                                             {TB, TB, WB}, {1.0 - 2.0 * TB, TB, WB}, {TB, 1.0 - 2.0 * TB, WB}};

// shape.rs:23-30
inline void barycentrics(const double n[3], double l[4]) {
    l[0] = 1.0 - n[0] - n[1] - n[2];
    l[1] = n[0];
    l[2] = n[1];
    l[3] = n[2];
}

// shape.rs:33-43
inline void shape_values(const double natural[3], double out[10]) {
    double l[4];
    barycentrics(natural, l);
    for (int i = 0; i < 4; ++i) out[i] = l[i] * (2.0 * l[i] - 1.0);
    for (int s = 0; s < 6; ++s) out[4 + s] = 4.0 * l[EDGE_CORNERS[s][0]] * l[EDGE_CORNERS[s][1]];
}

// shape.rs:46-67
inline void shape_derivatives(const double natural[3], double d[10][3]) {
    static const double DL[4][3] = {{-1.0, -1.0, -1.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    double l[4];
    barycentrics(natural, l);
    for (int axis = 0; axis < 3; ++axis) {
        for (int i = 0; i < 4; ++i) d[i][axis] = (4.0 * l[i] - 1.0) * DL[i][axis];
        for (int s = 0; s < 6; ++s) {
            int a = EDGE_CORNERS[s][0], b = EDGE_CORNERS[s][1];
            d[4 + s][axis] = 4.0 * (DL[a][axis] * l[b] + l[a] * DL[b][axis]);
        }
    }
}

struct Material {
    double e, nu;
    bool valid() const { return std::isfinite(e) && e > 0.0 && nu > -1.0 && nu < 0.5; }
    // material.rs:24-31
    void lame(double &lambda, double &mu) const {
        lambda = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
        mu = e / (2.0 * (1.0 + nu));
    }
};

struct PointGeometry {
    double gradients[10][3];
    double det_j;
    double position[3];
};

// element.rs:25-80; returns false on an inverted point
inline bool point_geometry(const Vec3 nodes[10], const double natural[3], PointGeometry &g) {
    double d[10][3], values[10];
    shape_derivatives(natural, d);
    shape_values(natural, values);
    double j[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double position[3] = {0, 0, 0};
    for (int a = 0; a < 10; ++a)
        for (int i = 0; i < 3; ++i) {
            for (int k = 0; k < 3; ++k) j[i][k] += d[a][i] * nodes[a][k];
            position[i] += values[a] * nodes[a][i];
        }
    double cof[3][3] = {
        {j[1][1] * j[2][2] - j[1][2] * j[2][1], j[0][2] * j[2][1] - j[0][1] * j[2][2], j[0][1] * j[1][2] - j[0][2] * j[1][1]},
        {j[1][2] * j[2][0] - j[1][0] * j[2][2], j[0][0] * j[2][2] - j[0][2] * j[2][0], j[0][2] * j[1][0] - j[0][0] * j[1][2]},
        {j[1][0] * j[2][1] - j[1][1] * j[2][0], j[0][1] * j[2][0] - j[0][0] * j[2][1], j[0][0] * j[1][1] - j[0][1] * j[1][0]}};
    double det_j = j[0][0] * cof[0][0] + j[0][1] * cof[1][0] + j[0][2] * cof[2][0];
    if (std::isnan(det_j) || det_j <= 0.0) return false;
    for (int a = 0; a < 10; ++a)
        for (int k = 0; k < 3; ++k) {
            double sum = cof[k][0] * d[a][0] + cof[k][1] * d[a][1] + cof[k][2] * d[a][2];
            g.gradients[a][k] = sum / det_j;
        }
    g.det_j = det_j;
    for (int i = 0; i < 3; ++i) g.position[i] = position[i];
    return true;
}

// ---------------------------------------------------------------------------------------------
// CSR: crates/solidfem-solve/src/csr.rs
// ---------------------------------------------------------------------------------------------

struct CsrMatrix {
    size_t size = 0;
    std::vector<size_t> row_starts;
    std::vector<u32> columns;
    std::vector<double> values;

    // csr.rs:19-71
    static CsrMatrix from_block_connectivity(size_t block_count, size_t block_size,
                                             const std::vector<std::vector<u32>> &elements) {
        std::vector<u64> pairs;
        for (const auto &el : elements)
            for (u32 row : el)
                for (u32 col : el) pairs.push_back(((u64)row << 32) | (u64)col);
        std::sort(pairs.begin(), pairs.end());
        pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
        CsrMatrix m;
        m.size = block_count * block_size;
        std::vector<size_t> block_rows(block_count + 1, 0);
        for (u64 p : pairs) block_rows[(p >> 32) + 1] += 1;
        for (size_t i = 0; i < block_count; ++i) block_rows[i + 1] += block_rows[i];
        m.row_starts.reserve(m.size + 1);
        m.columns.reserve(pairs.size() * block_size * block_size);
        m.row_starts.push_back(0);
        for (size_t b = 0; b < block_count; ++b)
            for (size_t r = 0; r < block_size; ++r) {
                for (size_t at = block_rows[b]; at < block_rows[b + 1]; ++at) {
                    size_t cb = (size_t)(pairs[at] & 0xffffffffu);
                    for (size_t o = 0; o < block_size; ++o) m.columns.push_back((u32)(cb * block_size + o));
                }
                m.row_starts.push_back(m.columns.size());
            }
        m.values.assign(m.columns.size(), 0.0);
        return m;
    }

    // csr.rs:89-96
    void add(size_t row, u32 column, double value) {
        auto b = columns.begin() + row_starts[row], e = columns.begin() + row_starts[row + 1];
        auto it = std::lower_bound(b, e, column);
        if (it == e || *it != column) throw std::logic_error("assembly writes only inside the pattern");
        values[it - columns.begin()] += value;
    }

    // csr.rs:99-105
    double get(size_t row, u32 column) const {
        auto b = columns.begin() + row_starts[row], e = columns.begin() + row_starts[row + 1];
        auto it = std::lower_bound(b, e, column);
        if (it == e || *it != column) return 0.0;
        return values[it - columns.begin()];
    }

    // csr.rs:112-132: rows [first, last) of A x, written at their own indices of y
    void multiply_rows(size_t first, size_t last, const double *x, double *y) const {
        for (size_t row = first; row < last; ++row) {
            double sum = 0.0;
            for (size_t at = row_starts[row]; at < row_starts[row + 1]; ++at) sum += values[at] * x[columns[at]];
            y[row] = sum;
        }
    }
    void multiply(const double *x, double *y) const { multiply_rows(0, size, x, y); }

    // csr.rs:142-154
    double multiply_rounding_bound(const double *x) const {
        const double eps = 2.220446049250313e-16;
        double sum = 0.0;
        for (size_t row = 0; row < size; ++row) {
            size_t s = row_starts[row], e = row_starts[row + 1];
            double row_sum = 0.0;
            for (size_t at = s; at < e; ++at) row_sum += std::fabs(values[at] * x[columns[at]]);
            double bound = (double)(e - s) * eps * row_sum;
            sum += bound * bound;
        }
        return std::sqrt(sum);
    }
};

}  // namespace refcpu
