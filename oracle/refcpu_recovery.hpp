// TEST INFRASTRUCTURE -- not part of the product. The oracle's restatement of stress recovery
// after a solve, in the reference's order of operations, so the device recovery can be checked
// bit for bit: integration-point stresses, the worst incident value at each node, the linear
// least-squares patch fit per mesh vertex, and the nodal field.
#pragma once

#include <cstring>
#include <limits>

#include "refcpu_core.hpp"

namespace refcpu {

// crates/solidfem-elastic/src/shape.rs:9-20
static const double NODE_NATURAL[10][3] = {{0.0, 0.0, 0.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0},
                                           {0.5, 0.0, 0.0}, {0.5, 0.5, 0.0}, {0.0, 0.5, 0.0}, {0.0, 0.0, 0.5},
                                           {0.5, 0.0, 0.5}, {0.0, 0.5, 0.5}};

// element.rs:135-146: h[i][j] = du_i/dx_j, accumulated over the nodes in order
inline void displacement_gradient(const PointGeometry &g, const double u[10][3], double h[3][3]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) h[i][j] = 0.0;
    for (int a = 0; a < 10; ++a)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) h[i][j] += u[a][i] * g.gradients[a][j];
}

// element.rs:160-171: [xx, yy, zz, 2xy, 2yz, 2zx]
inline void strain_at(const PointGeometry &g, const double u[10][3], double e[6]) {
    double h[3][3];
    displacement_gradient(g, u, h);
    e[0] = h[0][0];
    e[1] = h[1][1];
    e[2] = h[2][2];
    e[3] = h[0][1] + h[1][0];
    e[4] = h[1][2] + h[2][1];
    e[5] = h[2][0] + h[0][2];
}

// element.rs:174-186: [xx, yy, zz, xy, yz, zx]
inline void stress_at(const Material &m, const double e[6], double s[6]) {
    double lambda, mu;
    m.lame(lambda, mu);
    const double trace = e[0] + e[1] + e[2];
    s[0] = lambda * trace + 2.0 * mu * e[0];
    s[1] = lambda * trace + 2.0 * mu * e[1];
    s[2] = lambda * trace + 2.0 * mu * e[2];
    s[3] = mu * e[3];
    s[4] = mu * e[4];
    s[5] = mu * e[5];
}

// element.rs:189-192
inline double von_mises(const double s[6]) {
    const double a = s[0] - s[1], b = s[1] - s[2], c = s[2] - s[0];
    return std::sqrt(0.5 * (a * a + b * b + c * c) + 3.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
}

// patch.rs:9-80: a linear fit of each stress component about a vertex
struct PatchFit {
    double centre[3];
    double scale;
    double coefficients[4][6];

    static void offsets(const double p[3], const double c[3], double scale, double t[4]) {
        t[0] = 1.0;
        t[1] = (p[0] - c[0]) / scale;
        t[2] = (p[1] - c[1]) / scale;
        t[3] = (p[2] - c[2]) / scale;
    }

    // patch.rs:80-119: Gaussian elimination with partial pivoting
    static bool solve(double m[4][4], double rhs[4][6]) {
        const double SINGULAR_PIVOT = 1e-12;
        for (int column = 0; column < 4; ++column) {
            int pivot = column;
            for (int row = column + 1; row < 4; ++row)
                if (std::fabs(m[row][column]) > std::fabs(m[pivot][column])) pivot = row;
            if (std::fabs(m[pivot][column]) <= SINGULAR_PIVOT) return false;
            for (int k = 0; k < 4; ++k) std::swap(m[column][k], m[pivot][k]);
            for (int k = 0; k < 6; ++k) std::swap(rhs[column][k], rhs[pivot][k]);
            double pivot_row[4], pivot_rhs[6];
            for (int k = 0; k < 4; ++k) pivot_row[k] = m[column][k];
            for (int k = 0; k < 6; ++k) pivot_rhs[k] = rhs[column][k];
            for (int row = column + 1; row < 4; ++row) {
                const double factor = m[row][column] / pivot_row[column];
                for (int k = column; k < 4; ++k) m[row][k] -= factor * pivot_row[k];
                for (int k = 0; k < 6; ++k) rhs[row][k] -= factor * pivot_rhs[k];
            }
        }
        for (int column = 3; column >= 0; --column) {
            double value[6];
            for (int k = 0; k < 6; ++k) value[k] = rhs[column][k];
            for (int j = column + 1; j < 4; ++j)
                for (int k = 0; k < 6; ++k) value[k] -= m[column][j] * rhs[j][k];
            for (int k = 0; k < 6; ++k) rhs[column][k] = value[k] / m[column][column];
        }
        return true;
    }

    // patch.rs:35-64: samples are (position, stress) pairs in the order given
    static bool fit(const double centre[3], const std::vector<std::pair<Vec3, std::array<double, 6>>> &samples, PatchFit &out) {
        double scale = 0.0;
        for (const auto &s : samples)
            for (int axis = 0; axis < 3; ++axis) scale = std::fmax(scale, std::fabs(s.first[axis] - centre[axis]));
        if (samples.empty() || scale == 0.0) return false;
        double m[4][4] = {}, rhs[4][6] = {};
        for (const auto &s : samples) {
            double t[4];
            offsets(s.first.v, centre, scale, t);
            for (int i = 0; i < 4; ++i) {
                for (int j = 0; j < 4; ++j) m[i][j] += t[i] * t[j];
                for (int c = 0; c < 6; ++c) rhs[i][c] += t[i] * s.second[c];
            }
        }
        if (!solve(m, rhs)) return false;
        for (int a = 0; a < 3; ++a) out.centre[a] = centre[a];
        out.scale = scale;
        for (int i = 0; i < 4; ++i)
            for (int c = 0; c < 6; ++c) out.coefficients[i][c] = rhs[i][c];
        return true;
    }

    // patch.rs:67-78
    void at(const double p[3], double value[6]) const {
        double t[4];
        offsets(p, centre, scale, t);
        for (int c = 0; c < 6; ++c) value[c] = 0.0;
        for (int i = 0; i < 4; ++i)
            for (int c = 0; c < 6; ++c) value[c] += t[i] * coefficients[i][c];
    }
};

// results.rs:555-571: nearest integration point to a node in natural coordinates, first on a tie
inline int nearest_natural_point(const QPoint *points, int count, const double node[3]) {
    double best = std::numeric_limits<double>::max();
    int index = 0;
    for (int p = 0; p < count; ++p) {
        double d = 0.0;
        for (int a = 0; a < 3; ++a) d += (points[p].natural[a] - node[a]) * (points[p].natural[a] - node[a]);
        if (d < best) {
            best = d;
            index = p;
        }
    }
    return index;
}

struct RecoveredField {
    std::vector<double> point_position;  // 3 per (element, point)
    std::vector<double> point_stress;    // 6 per (element, point)
    std::vector<double> point_von_mises;
    std::vector<double> nodal_stress;    // 6 per node
    std::vector<double> nodal_von_mises;
    int64_t fault_tet = -1, fault_point = -1;  // an inverted element, when the recovery failed
    std::vector<PatchFit> fits;                // per vertex, meaningful where fit_state is FITTED
    std::vector<char> fit_state;               // per vertex: UNFITTED (flat), FITTED or MIXED
};
enum : char { UNFITTED = 0, FITTED = 1, MIXED = 2 };

// Recovery::advance and Recovery::into_field (results.rs:376-536) for nodal displacements `u`
// (3 per node); the material comes from `field` at each point.
template <typename Field>
inline RecoveredField recover(const QuadMesh &mesh, const Field &field, int rule, const std::vector<double> &u) {
    RecoveredField out;
    const QPoint *points = rule == 14 ? FOURTEEN_POINT : FOUR_POINT;
    const int count = rule == 14 ? 14 : 4;
    int nearest[10];
    for (int n = 0; n < 10; ++n) nearest[n] = nearest_natural_point(points, count, NODE_NATURAL[n]);
    const size_t elements = mesh.tets.size(), nodes = mesh.nodes.size(), vertices = mesh.vertex_count;
    std::vector<Material> point_material;
    point_material.reserve(elements * count);
    std::vector<std::array<double, 6>> worst_stress(nodes, std::array<double, 6>{0, 0, 0, 0, 0, 0});
    std::vector<double> worst_vm(nodes, std::numeric_limits<double>::lowest());
    for (size_t t = 0; t < elements; ++t) {
        const auto &tet = mesh.tets[t];
        Vec3 en[10];
        double ue[10][3];
        for (int a = 0; a < 10; ++a) {
            en[a] = mesh.nodes[tet[a]];
            for (int i = 0; i < 3; ++i) ue[a][i] = u[3 * (size_t)tet[a] + i];
        }
        for (int p = 0; p < count; ++p) {
            PointGeometry g;
            Material m{};
            if (!point_geometry(en, points[p].natural, g) || !(m = field.at(g.position), m.valid())) {
                out.fault_tet = (int64_t)t;
                out.fault_point = p;
                return out;
            }
            double e[6], s[6];
            strain_at(g, ue, e);
            stress_at(m, e, s);
            point_material.push_back(m);
            for (int a = 0; a < 3; ++a) out.point_position.push_back(g.position[a]);
            for (int c = 0; c < 6; ++c) out.point_stress.push_back(s[c]);
            out.point_von_mises.push_back(von_mises(s));
        }
        const Material *own = &point_material[point_material.size() - (size_t)count];
        for (int n = 0; n < 10; ++n) {
            PointGeometry g;
            if (!point_geometry(en, NODE_NATURAL[n], g)) {
                out.fault_tet = (int64_t)t;
                out.fault_point = n;
                return out;
            }
            double e[6], s[6];
            strain_at(g, ue, e);
            stress_at(own[nearest[n]], e, s);
            const double vm = von_mises(s);
            if (vm > worst_vm[tet[n]]) {
                worst_vm[tet[n]] = vm;
                for (int c = 0; c < 6; ++c) worst_stress[tet[n]][c] = s[c];
            }
        }
    }
    // build_patches, results.rs:414-436: the elements at each vertex, in element order
    std::vector<size_t> offsets(vertices + 1, 0);
    for (const auto &tet : mesh.tets)
        for (int c = 0; c < 4; ++c) offsets[tet[c] + 1] += 1;
    for (size_t v = 1; v <= vertices; ++v) offsets[v] += offsets[v - 1];
    std::vector<size_t> next(offsets.begin(), offsets.end() - 1);
    std::vector<u32> incident(offsets[vertices]);
    for (size_t t = 0; t < elements; ++t)
        for (int c = 0; c < 4; ++c) incident[next[mesh.tets[t][c]]++] = (u32)t;
    // fit_next_vertex, results.rs:438-467
    std::vector<PatchFit> fits(vertices);
    std::vector<char> fitted(vertices, 0);
    for (size_t v = 0; v < vertices; ++v) {
        std::vector<std::pair<Vec3, std::array<double, 6>>> samples;
        bool uniform = true;
        const Material *first = nullptr;
        for (size_t at = offsets[v]; at < offsets[v + 1]; ++at) {
            const size_t start = (size_t)incident[at] * count;
            for (size_t p = start; p < start + (size_t)count; ++p) {
                const Material &m = point_material[p];
                if (!first) first = &m;
                else if (std::memcmp(&first->e, &m.e, 8) != 0 || std::memcmp(&first->nu, &m.nu, 8) != 0) uniform = false;
                Vec3 pos{out.point_position[3 * p], out.point_position[3 * p + 1], out.point_position[3 * p + 2]};
                std::array<double, 6> s;
                for (int c = 0; c < 6; ++c) s[c] = out.point_stress[6 * p + c];
                samples.push_back({pos, s});
            }
        }
        if (uniform) fitted[v] = PatchFit::fit(mesh.nodes[v].v, samples, fits[v]) ? 1 : 0;
        else fitted[v] = 2;
    }
    out.fits = fits;
    out.fit_state.resize(vertices);
    for (size_t v = 0; v < vertices; ++v) out.fit_state[v] = fitted[v] == 1 ? FITTED : fitted[v] == 2 ? MIXED : UNFITTED;
    // into_field, results.rs:376-411
    out.nodal_stress.assign(6 * nodes, 0.0);
    for (size_t v = 0; v < vertices; ++v) {
        double s[6];
        if (fitted[v] == 1) fits[v].at(mesh.nodes[v].v, s);
        else
            for (int c = 0; c < 6; ++c) s[c] = worst_stress[v][c];
        for (int c = 0; c < 6; ++c) out.nodal_stress[6 * v + c] = s[c];
    }
    std::vector<char> done(nodes > vertices ? nodes - vertices : 0, 0);
    for (const auto &tet : mesh.tets)
        for (int slot = 0; slot < 6; ++slot) {
            const u32 midside = tet[4 + slot];
            if (done[midside - vertices]) continue;
            done[midside - vertices] = 1;
            const u32 a = tet[EDGE_CORNERS[slot][0]], b = tet[EDGE_CORNERS[slot][1]];
            double s[6];
            if (fitted[a] == 1 && fitted[b] == 1) {
                // results.rs:544-552
                double p[6], q[6];
                fits[a].at(mesh.nodes[midside].v, p);
                fits[b].at(mesh.nodes[midside].v, q);
                for (int c = 0; c < 6; ++c) s[c] = 0.5 * (p[c] + q[c]);
            } else {
                for (int c = 0; c < 6; ++c) s[c] = worst_stress[midside][c];
            }
            for (int c = 0; c < 6; ++c) out.nodal_stress[6 * (size_t)midside + c] = s[c];
        }
    out.nodal_von_mises.resize(nodes);
    for (size_t n = 0; n < nodes; ++n) out.nodal_von_mises[n] = von_mises(&out.nodal_stress[6 * n]);
    return out;
}

// The refusals of an error estimate, numbered as the product's status codes.
enum EstimateStatus {
    ESTIMATE_OK = 0,
    ESTIMATE_INVERTED = 45,     // a = element, b = point of the 14-point rule
    ESTIMATE_MATERIAL = 46,     // a = element, b = point
    ESTIMATE_MIXED = 53,        // a = vertex
    ESTIMATE_COPLANAR = 54,     // a = vertex
    ESTIMATE_ZERO_ENERGY = 55,
    ESTIMATE_NON_FINITE = 56,
};

struct ErrorEstimate {
    std::vector<double> element_errors;  // e_K, in element order
    double error_sum = 0.0, energy_sum = 0.0, relative_error = 0.0;
    int status = ESTIMATE_OK;
    int64_t a = -1, b = -1;
};

// compliance_energy, error_estimate.rs:240-248: s : C^-1 : s for an isotropic material
inline double compliance_energy(const Material &m, const double s[6]) {
    const double nu = m.nu;
    const double normal = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
    const double coupling = s[0] * s[1] + s[1] * s[2] + s[2] * s[0];
    const double shear = s[3] * s[3] + s[4] * s[4] + s[5] * s[5];
    return (normal - 2.0 * nu * coupling + 2.0 * (1.0 + nu) * shear) / m.e;
}

// ErrorEstimation::advance over every element, then ErrorEstimation::estimate
// (error_estimate.rs:121-171): per element, element_integrals (:174-205) with the 14-point rule
// whatever rule the solve used, the recovered stress at the element's nodes from the patch fits
// (recovered_at_nodes, :207-231: a corner's fit at the corner, a midside the mean of its edge's
// two fits at the midside), and the running sums in element order.
template <typename Field>
inline ErrorEstimate estimate_error(const QuadMesh &mesh, const Field &field, const std::vector<double> &u,
                                    const RecoveredField &r) {
    ErrorEstimate out;
    for (size_t t = 0; t < mesh.tets.size(); ++t) {
        const auto &tet = mesh.tets[t];
        for (int corner = 0; corner < 4; ++corner) {
            const u32 v = tet[corner];
            if (r.fit_state[v] != FITTED) {
                out.status = r.fit_state[v] == MIXED ? ESTIMATE_MIXED : ESTIMATE_COPLANAR;
                out.a = v;
                return out;
            }
        }
        double recovered[10][6];
        for (int corner = 0; corner < 4; ++corner) r.fits[tet[corner]].at(mesh.nodes[tet[corner]].v, recovered[corner]);
        for (int slot = 0; slot < 6; ++slot) {
            const double *at = mesh.nodes[tet[4 + slot]].v;
            double p[6], q[6];
            r.fits[tet[EDGE_CORNERS[slot][0]]].at(at, p);
            r.fits[tet[EDGE_CORNERS[slot][1]]].at(at, q);
            for (int k = 0; k < 6; ++k) recovered[4 + slot][k] = 0.5 * (p[k] + q[k]);
        }
        Vec3 en[10];
        double ue[10][3];
        for (int a = 0; a < 10; ++a) {
            en[a] = mesh.nodes[tet[a]];
            for (int i = 0; i < 3; ++i) ue[a][i] = u[3 * (size_t)tet[a] + i];
        }
        double error = 0.0, energy = 0.0;
        for (int q = 0; q < 14; ++q) {
            const QPoint &point = FOURTEEN_POINT[q];
            PointGeometry g;
            if (!point_geometry(en, point.natural, g)) {
                out.status = ESTIMATE_INVERTED;
                out.a = (int64_t)t;
                out.b = q;
                return out;
            }
            const Material m = field.at(g.position);
            if (!m.valid()) {
                out.status = ESTIMATE_MATERIAL;
                out.a = (int64_t)t;
                out.b = q;
                return out;
            }
            double e[6], stress[6], shape[10], difference[6] = {0, 0, 0, 0, 0, 0};
            strain_at(g, ue, e);
            stress_at(m, e, stress);
            shape_values(point.natural, shape);
            for (int a = 0; a < 10; ++a)
                for (int k = 0; k < 6; ++k) difference[k] += shape[a] * recovered[a][k];
            for (int k = 0; k < 6; ++k) difference[k] -= stress[k];
            const double measure = point.weight * g.det_j / 6.0;
            error += measure * compliance_energy(m, difference);
            energy += measure * compliance_energy(m, stress);
        }
        out.element_errors.push_back(std::sqrt(error));
        out.error_sum += error;
        out.energy_sum += energy;
    }
    if (!std::isfinite(out.energy_sum) || !std::isfinite(out.error_sum)) {
        out.status = ESTIMATE_NON_FINITE;
        return out;
    }
    if (out.energy_sum <= 0.0) {
        out.status = ESTIMATE_ZERO_ENERGY;
        return out;
    }
    out.relative_error = std::sqrt(out.error_sum / (out.energy_sum + out.error_sum));
    return out;
}

// Peak (peak_with_holder, results.rs:184-226): the candidates are the nodes, then the points, that
// `node_inside` / `point_inside` keep (the caller's region test, evaluated by the caller); the
// fold keeps the peak when it is NaN, else takes a candidate that is NaN or strictly larger.
struct Peak {
    bool found = false, at_point = false;
    size_t index = 0;
    double von_mises = 0.0;
    double position[3] = {0.0, 0.0, 0.0};
};

inline Peak peak_of(const double *positions, size_t nodes, const double *nodal_vm, size_t stride, const double *points,
                    size_t count, const unsigned char *node_inside, const unsigned char *point_inside) {
    Peak peak;
    auto fold = [&](bool at_point, size_t index, double value, const double *position) {
        if (peak.found) {
            if (std::isnan(peak.von_mises)) return;
            if (!(std::isnan(value) || value > peak.von_mises)) return;
        }
        peak.found = true;
        peak.at_point = at_point;
        peak.index = index;
        peak.von_mises = value;
        for (int a = 0; a < 3; ++a) peak.position[a] = position[a];
    };
    for (size_t n = 0; n < nodes; ++n)
        if (!node_inside || node_inside[n]) fold(false, n, nodal_vm[n * stride], positions + 3 * n);
    for (size_t k = 0; k < count; ++k)
        if (!point_inside || point_inside[k]) fold(true, k, points[10 * k + 9], points + 10 * k);
    return peak;
}

// nodal_rotations (results.rs:262-305): per element in order, per node in order, the node's
// geometry, rotation_at (element.rs:149-156) and the determinant-weighted sums; then each node's
// sum over its weight. Returns the first element whose geometry fails at a node, else -1.
inline int64_t nodal_rotations(const QuadMesh &mesh, const std::vector<double> &u, std::vector<double> &out) {
    const size_t nodes = mesh.nodes.size();
    std::vector<double> weighted(3 * nodes, 0.0), weights(nodes, 0.0);
    for (size_t t = 0; t < mesh.tets.size(); ++t) {
        const auto &tet = mesh.tets[t];
        Vec3 en[10];
        double ue[10][3];
        for (int a = 0; a < 10; ++a) {
            en[a] = mesh.nodes[tet[a]];
            for (int i = 0; i < 3; ++i) ue[a][i] = u[3 * (size_t)tet[a] + i];
        }
        for (int n = 0; n < 10; ++n) {
            PointGeometry g;
            if (!point_geometry(en, NODE_NATURAL[n], g)) return (int64_t)t;
            double h[3][3];
            displacement_gradient(g, ue, h);
            const double rotation[3] = {0.5 * (h[2][1] - h[1][2]), 0.5 * (h[0][2] - h[2][0]), 0.5 * (h[1][0] - h[0][1])};
            const size_t global = tet[n];
            for (int axis = 0; axis < 3; ++axis) weighted[3 * global + axis] += g.det_j * rotation[axis];
            weights[global] += g.det_j;
        }
    }
    out.resize(3 * nodes);
    for (size_t n = 0; n < nodes; ++n)
        for (int axis = 0; axis < 3; ++axis) out[3 * n + axis] = weighted[3 * n + axis] / weights[n];
    return -1;
}

// MeshComparison::between (comparison.rs:101-141) of a coarse and a fine solve, each given as its
// mesh, nodal von Mises (stride apart), point values (10 per point, element-major) and the masks of
// the candidates inside the region (evaluated by the caller; NULL: all).
struct ComparedSide {
    const QuadMesh *mesh;
    const double *nodal_vm;
    size_t nodes, stride;
    const double *points;
    size_t count;
    const unsigned char *node_inside, *point_inside;
};
struct ComparedPeak {
    Peak peak;
    size_t elements = 0;
    double volume = 0.0;
};
struct Comparison {
    int status = 0;  // 0, or the refusal as the product's status code
    int64_t a = -1, b = -1;
    double value = 0.0;
    ComparedPeak coarse, fine;
    double change = 0.0, distance = 0.0;
};

inline Comparison compare_meshes(const ComparedSide &coarse, const ComparedSide &fine) {
    Comparison out;
    auto refuse = [&](int status, int64_t a, int64_t b = -1, double value = 0.0) {
        out.status = status;
        out.a = a;
        out.b = b;
        out.value = value;
        return out;
    };
    const ComparedSide *sides[2] = {&coarse, &fine};
    // check_matches, comparison.rs:199-224
    for (int w = 0; w < 2; ++w) {
        const ComparedSide &s = *sides[w];
        const size_t tets = s.mesh->tets.size();
        if (s.nodes != s.mesh->nodes.size() || tets == 0 || s.count == 0 || s.count % tets != 0) return refuse(58, w);
    }
    // region_peak, comparison.rs:271-305
    ComparedPeak *peaks[2] = {&out.coarse, &out.fine};
    for (int w = 0; w < 2; ++w) {
        const ComparedSide &s = *sides[w];
        bool any = false;
        for (size_t k = 0; k < s.count && !any; ++k) any = !s.point_inside || s.point_inside[k];
        if (!any) return refuse(59, w);
        std::vector<double> positions(3 * s.mesh->nodes.size());
        for (size_t n = 0; n < s.mesh->nodes.size(); ++n)
            for (int a = 0; a < 3; ++a) positions[3 * n + a] = s.mesh->nodes[n][a];
        const Peak peak = peak_of(positions.data(), s.nodes, s.nodal_vm, s.stride, s.points, s.count, s.node_inside, s.point_inside);
        if (!peak.found) return refuse(59, w);
        if (!std::isfinite(peak.von_mises)) return refuse(70, w, -1, peak.von_mises);
        if (peak.von_mises <= 0.0) return refuse(71, w);
        peaks[w]->peak = peak;
    }
    // elements_inside, comparison.rs:226-239
    for (int w = 0; w < 2; ++w) {
        const ComparedSide &s = *sides[w];
        const size_t per = s.count / s.mesh->tets.size();
        int64_t last = -1;
        for (size_t k = 0; k < s.count; ++k) {
            const int64_t tet = (int64_t)(k / per);
            if (last != tet && (!s.point_inside || s.point_inside[k])) {
                peaks[w]->elements += 1;
                last = tet;
            }
        }
    }
    if (out.fine.elements <= out.coarse.elements) return refuse(72, (int64_t)out.coarse.elements, (int64_t)out.fine.elements);
    out.change = (out.fine.peak.von_mises - out.coarse.peak.von_mises) / out.fine.peak.von_mises;
    double squared = 0.0;
    for (int a = 0; a < 3; ++a) {
        const double d = out.fine.peak.position[a] - out.coarse.peak.position[a];
        squared += d * d;
    }
    out.distance = std::sqrt(squared);
    // holder_volume, comparison.rs:241-269: the corner tetrahedron's volume (tet.rs:9-16)
    for (int w = 0; w < 2; ++w) {
        const ComparedSide &s = *sides[w];
        const QuadMesh &m = *s.mesh;
        auto volume = [&](const std::array<u32, 10> &t) { return tet_volume(m.nodes[t[0]], m.nodes[t[1]], m.nodes[t[2]], m.nodes[t[3]]); };
        const Peak &peak = peaks[w]->peak;
        if (peak.at_point) {
            peaks[w]->volume = volume(m.tets[peak.index / (s.count / m.tets.size())]);
        } else {
            double sum = 0.0;
            size_t n = 0;
            for (const auto &t : m.tets)
                for (int a = 0; a < 10; ++a)
                    if (t[a] == (u32)peak.index) {
                        sum += volume(t);
                        n += 1;
                        break;
                    }
            peaks[w]->volume = sum / (double)n;
        }
    }
    return out;
}

}  // namespace refcpu
