// refcpu: TEST INFRASTRUCTURE (the oracle) -- the study, the conjugate gradient solve and the
// C ABI the tests and the CPU-baseline leg of the benchmark load. Restates
// crates/solidfem-elastic/src/{assembly,loads,rigid_modes,study}.rs and
// crates/solidfem-solve/src/{pcg,preconditioner}.rs in the reference's floating-point order.
#include "refcpu.h"

#include <chrono>
#include <functional>
#include <map>
#include <memory>

#include "refcpu_multigrid.hpp"
#include "refcpu_recovery.hpp"

namespace refcpu {

static double now_seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------------------------
// Workload inputs (seeded; the same definitions the product's mesh generator uses)
// ---------------------------------------------------------------------------------------------

inline u64 splitmix64(u64 x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
inline double unit_uniform(u64 bits) { return (double)(bits >> 11) * (1.0 / 9007199254740992.0); }

// Jitter of interior vertices: offset = (2u - 1) * (fraction * h), u = uniform(seed ^ (3 v + axis)).
static void jitter_vertices(LinearMesh &m, const ref_workload &w) {
    if (!(w.jitter_fraction > 0.0)) return;
    size_t d0 = (size_t)w.divisions[0] + 1, d1 = (size_t)w.divisions[1] + 1;
    double h[3];
    for (int a = 0; a < 3; ++a) h[a] = w.size[a] / (double)w.divisions[a];
    for (size_t v = 0; v < m.positions.size(); ++v) {
        size_t i = v % d0, j = (v / d0) % d1, k = v / (d0 * d1);
        bool interior = i > 0 && i < (size_t)w.divisions[0] && j > 0 && j < (size_t)w.divisions[1] && k > 0 &&
                        k < (size_t)w.divisions[2];
        if (!interior) continue;
        for (int a = 0; a < 3; ++a) {
            double u = unit_uniform(splitmix64(w.jitter_seed ^ (u64)(3 * v + a)));
            m.positions[v][a] = m.positions[v][a] + (2.0 * u - 1.0) * (w.jitter_fraction * h[a]);
        }
    }
}

// N(0, 1) by Box-Muller from two counter-based uniforms.
static double normal_sample(u64 seed, u64 counter) {
    double u1 = ((double)(splitmix64(seed + 2 * counter) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    double u2 = unit_uniform(splitmix64(seed + 2 * counter + 1));
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// ---------------------------------------------------------------------------------------------
// Model, load, assembly: elastic/src/{study,loads,assembly,rigid_modes}.rs
// ---------------------------------------------------------------------------------------------

struct Traction {
    BoundaryFace face;
    double t[3];
};

struct MaterialField {
    int kind = REF_MATERIAL_UNIFORM;
    Material shell{1, 0}, core{1, 0};
    double depth = 0;
    double lo[3], hi[3];
    int transverse[2] = {1, 2};
    Material at(const double p[3]) const {
        if (kind == REF_MATERIAL_UNIFORM) return shell;
        double d = 1e300;
        for (int s = 0; s < 2; ++s) {
            int a = transverse[s];
            d = std::min(d, std::min(p[a] - lo[a], hi[a] - p[a]));
        }
        return d < depth ? shell : core;
    }
};

struct Problem {
    QuadMesh mesh;
    std::vector<BoundaryFace> boundary;
    std::vector<char> is_fixed;
    std::vector<Traction> tractions;
    MaterialField material;
    int quadrature = 4;
    std::vector<u32> free;
    size_t free_count = 0;
    std::vector<double> rhs;
    CsrMatrix matrix;
    double assembly_seconds = 0;
    // Model::prescribe_displacement (study.rs:90-99): the map keyed by (node, axis), the last call
    // winning, then study.rs:661-683's (free dof, value) list sorted by dof
    std::map<std::pair<u32, u32>, double> prescribed_nodes;
    std::vector<std::pair<size_t, double>> prescribed;
};

// study.rs:661-683: each prescribed component as (free dof, value), sorted by dof; refuses an axis
// above 2, a non-finite value, a node out of range or a fixed node
static std::vector<std::pair<size_t, double>> prescribed_dofs(const Problem &p) {
    std::vector<std::pair<size_t, double>> dofs;
    for (const auto &kv : p.prescribed_nodes) {
        const u32 node = kv.first.first, axis = kv.first.second;
        if (axis > 2) throw std::runtime_error("axis out of range");
        if (!std::isfinite(kv.second)) throw std::runtime_error("non-finite prescribed value");
        if ((size_t)node >= p.free.size()) throw std::runtime_error("node out of range");
        if (p.free[node] == FIXED) throw std::runtime_error("prescribed on a fixed node");
        dofs.push_back({3 * (size_t)p.free[node] + axis, kv.second});
    }
    std::stable_sort(dofs.begin(), dofs.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    return dofs;
}

// study.rs:567-593: symmetric elimination. Each prescribed row's off-diagonal entries (read before
// this row's own updates) move `entry * value` off the right-hand side of the unconstrained rows
// and are cancelled in both the row and the column; then the prescribed rows' right-hand sides
// become diagonal * value.
static void apply_prescribed(Problem &p) {
    std::vector<char> constrained(p.rhs.size(), 0);
    for (const auto &pr : p.prescribed) constrained[pr.first] = 1;
    for (const auto &pr : p.prescribed) {
        const size_t dof = pr.first;
        const double value = pr.second;
        std::vector<std::pair<u32, double>> entries;
        for (size_t at = p.matrix.row_starts[dof]; at < p.matrix.row_starts[dof + 1]; ++at)
            if ((size_t)p.matrix.columns[at] != dof) entries.push_back({p.matrix.columns[at], p.matrix.values[at]});
        for (const auto &e : entries) {
            const size_t c = e.first;
            if (!constrained[c]) p.rhs[c] -= e.second * value;
            p.matrix.add(c, (u32)dof, -e.second);
            p.matrix.add(dof, e.first, -e.second);
        }
    }
    for (const auto &pr : p.prescribed) p.rhs[pr.first] = p.matrix.get(pr.first, (u32)pr.first) * pr.second;
}

// element.rs:98-133
static bool element_stiffness(const Vec3 nodes[10], const MaterialField &field, int rule, double ke[900]) {
    const QPoint *points = rule == 14 ? FOURTEEN_POINT : FOUR_POINT;
    int count = rule == 14 ? 14 : 4;
    for (int i = 0; i < 900; ++i) ke[i] = 0.0;
    for (int index = 0; index < count; ++index) {
        const QPoint &q = points[index];
        PointGeometry g;
        if (!point_geometry(nodes, q.natural, g)) return false;
        Material m = field.at(g.position);
        if (!m.valid()) return false;
        double lambda, mu;
        m.lame(lambda, mu);
        double measure = q.weight * g.det_j / 6.0;
        for (int a = 0; a < 10; ++a)
            for (int b = a; b < 10; ++b) {
                const double *ga = g.gradients[a], *gb = g.gradients[b];
                double dot = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
                for (int i = 0; i < 3; ++i)
                    for (int k = 0; k < 3; ++k) {
                        double value = lambda * (ga[i] * gb[k]) + mu * (ga[k] * gb[i]);
                        if (i == k) value += mu * dot;
                        value *= measure;
                        ke[(3 * a + i) * 30 + (3 * b + k)] += value;
                        if (a != b) ke[(3 * b + k) * 30 + (3 * a + i)] += value;
                    }
            }
    }
    return true;
}

// loads.rs:44-71
static void triangle_shape(double r, double s, double values[6], double dr[6], double ds[6]) {
    double l1 = 1.0 - r - s;
    values[0] = l1 * (2.0 * l1 - 1.0);
    values[1] = r * (2.0 * r - 1.0);
    values[2] = s * (2.0 * s - 1.0);
    values[3] = 4.0 * l1 * r;
    values[4] = 4.0 * r * s;
    values[5] = 4.0 * s * l1;
    dr[0] = -(4.0 * l1 - 1.0);
    dr[1] = 4.0 * r - 1.0;
    dr[2] = 0.0;
    dr[3] = 4.0 * (l1 - r);
    dr[4] = 4.0 * s;
    dr[5] = -4.0 * s;
    ds[0] = -(4.0 * l1 - 1.0);
    ds[1] = 0.0;
    ds[2] = 4.0 * s - 1.0;
    ds[3] = -4.0 * r;
    ds[4] = 4.0 * r;
    ds[5] = 4.0 * (l1 - s);
}

// loads.rs:10-41
static void add_traction(const QuadMesh &mesh, const BoundaryFace &face, const double traction[3], std::vector<double> &load) {
    Vec3 positions[6];
    for (int n = 0; n < 6; ++n) positions[n] = mesh.nodes[face.nodes[n]];
    for (int qp = 0; qp < 6; ++qp) {
        double r = TRIANGLE_POINTS[qp][0], s = TRIANGLE_POINTS[qp][1], weight = TRIANGLE_POINTS[qp][2];
        double values[6], dr[6], ds[6];
        triangle_shape(r, s, values, dr, ds);
        double tr[3] = {0, 0, 0}, ts[3] = {0, 0, 0};
        for (int n = 0; n < 6; ++n)
            for (int a = 0; a < 3; ++a) {
                tr[a] += dr[n] * positions[n][a];
                ts[a] += ds[n] * positions[n][a];
            }
        double normal[3] = {tr[1] * ts[2] - tr[2] * ts[1], tr[2] * ts[0] - tr[0] * ts[2], tr[0] * ts[1] - tr[1] * ts[0]};
        double length = std::sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
        double measure = 0.5 * length * weight;
        for (int n = 0; n < 6; ++n)
            for (int a = 0; a < 3; ++a) load[3 * (size_t)face.nodes[n] + a] += traction[a] * values[n] * measure;
    }
}

// rigid_modes.rs:36-83
static std::vector<double> free_rigid_body_modes(const std::vector<Vec3> &nodes, const std::vector<u32> &free, size_t free_count) {
    double center[3] = {0, 0, 0};
    for (size_t n = 0; n < free.size(); ++n)
        if (free[n] != FIXED)
            for (int a = 0; a < 3; ++a) center[a] += nodes[n][a];
    double divisor = (double)std::max<size_t>(free_count, 1);
    for (int a = 0; a < 3; ++a) center[a] /= divisor;
    double reach = 0.0;
    for (size_t n = 0; n < free.size(); ++n)
        if (free[n] != FIXED)
            for (int a = 0; a < 3; ++a) reach = std::max(reach, std::fabs(nodes[n][a] - center[a]));
    double scale = reach > 0.0 ? 1.0 / reach : 1.0;
    std::vector<double> modes(3 * free_count * 6, 0.0);
    for (size_t n = 0; n < free.size(); ++n) {
        if (free[n] == FIXED) continue;
        size_t at = 3 * (size_t)free[n] * 6;
        double x = (nodes[n][0] - center[0]) * scale;
        double y = (nodes[n][1] - center[1]) * scale;
        double z = (nodes[n][2] - center[2]) * scale;
        double rows[3][6] = {{1, 0, 0, 0, z, -y}, {0, 1, 0, -z, 0, x}, {0, 0, 1, y, -x, 0}};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 6; ++c) modes[at + r * 6 + c] = rows[r][c];
    }
    return modes;
}

static std::unique_ptr<Problem> build_problem(const ref_workload &w, const uint32_t *pnodes = nullptr, const uint32_t *paxes = nullptr,
                                              const double *pvalues = nullptr, int64_t pcount = 0) {
    double t0 = now_seconds();
    auto p = std::make_unique<Problem>();
    for (int64_t k = 0; k < pcount; ++k) p->prescribed_nodes[{pnodes[k], paxes[k]}] = pvalues[k];
    LinearMesh linear;
    if (w.wall_cells > 0) {
        // a bar along its longest axis with a hollow section (demo/wasm/src/part.rs:87-185)
        int along = 0;
        for (int a = 1; a < 3; ++a)
            if (w.size[a] > w.size[along]) along = a;
        const int across[2] = {(along + 1) % 3, (along + 2) % 3};
        std::vector<double> lines[3];
        for (int a = 0; a < 3; ++a) {
            if (a == along) {
                for (int i = 0; i <= w.divisions[a]; ++i) lines[a].push_back(box_coord(w.origin[a], w.size[a], (size_t)i, (size_t)w.divisions[a]));
            } else {
                for (double x : tube_cross_lines(w.size[a], w.wall_thickness, w.wall_cells, w.divisions[a] - 2 * w.wall_cells))
                    lines[a].push_back(w.origin[a] + x);
            }
        }
        linear = tube_tet_mesh(lines, across, w.wall_cells);
    } else {
        linear = box_tet_mesh(w.origin, w.size, w.divisions);
        jitter_vertices(linear, w);
    }
    validate_linear(linear);
    p->mesh = quadratic_from_linear(linear);
    p->boundary = boundary_faces(p->mesh);
    const auto &nodes = p->mesh.nodes;
    p->is_fixed.assign(nodes.size(), 0);
    if (w.clamp_axis >= 0)
        for (size_t n = 0; n < nodes.size(); ++n)
            if (std::fabs(nodes[n][w.clamp_axis] - w.clamp_value) < 1e-9) p->is_fixed[n] = 1;
    // The loaded faces: boundary faces whose corners lie on the plane (support.rs:80-89); none
    // when load_axis < 0.
    size_t face_index = 0;
    const bool loads = w.load_axis >= 0;
    int u = (w.load_axis + 1) % 3, v = (w.load_axis + 2) % 3;
    for (const auto &f : p->boundary) {
        if (!loads) break;
        bool on = true;
        for (int c = 0; c < 3; ++c)
            if (!(std::fabs(nodes[f.nodes[c]][w.load_axis] - w.load_value) < 1e-9)) on = false;
        if (!on) continue;
        Traction t;
        t.face = f;
        if (w.traction_kind == REF_TRACTION_UNIFORM) {
            for (int a = 0; a < 3; ++a) t.t[a] = w.traction[a];
        } else if (w.traction_kind == REF_TRACTION_RANDOM) {
            for (int a = 0; a < 3; ++a) {
                u64 counter = ((u64)(uint32_t)w.load_case << 40) ^ ((u64)face_index << 2) ^ (u64)a;
                t.t[a] = w.traction_scale * normal_sample(w.traction_seed, counter);
            }
        } else {
            const Vec3 &a = nodes[f.nodes[0]], &b = nodes[f.nodes[1]], &c = nodes[f.nodes[2]];
            double cu = (a[u] + b[u] + c[u]) / 3.0 - w.axis_u;
            double cv = (a[v] + b[v] + c[v]) / 3.0 - w.axis_v;
            double r = std::sqrt(cu * cu + cv * cv);
            t.t[w.load_axis] = 0.0;
            t.t[u] = -w.traction_scale * cv / r;
            t.t[v] = w.traction_scale * cu / r;
        }
        p->tractions.push_back(t);
        ++face_index;
    }
    // Material field
    p->material.kind = w.material_kind;
    p->material.shell = Material{w.youngs, w.poisson};
    p->material.core = Material{w.core_youngs, w.core_poisson};
    p->material.depth = w.shell_depth;
    int longest = 0;
    for (int a = 1; a < 3; ++a)
        if (w.size[a] > w.size[longest]) longest = a;
    p->material.transverse[0] = (longest + 1) % 3;
    p->material.transverse[1] = (longest + 2) % 3;
    for (int a = 0; a < 3; ++a) {
        p->material.lo[a] = w.origin[a];
        p->material.hi[a] = w.origin[a] + w.size[a];
    }
    p->quadrature = w.quadrature_points == 14 ? 14 : 4;

    // study.rs:419-474: load, free numbering, restricted right-hand side, pattern
    std::vector<double> load(3 * nodes.size(), 0.0);
    for (const auto &t : p->tractions) add_traction(p->mesh, t.face, t.t, load);
    p->free.assign(nodes.size(), FIXED);
    for (size_t n = 0; n < nodes.size(); ++n)
        if (!p->is_fixed[n]) p->free[n] = (u32)p->free_count++;
    p->rhs.reserve(3 * p->free_count);
    for (size_t n = 0; n < nodes.size(); ++n)
        if (p->free[n] != FIXED)
            for (int a = 0; a < 3; ++a) p->rhs.push_back(load[3 * n + a]);
    p->prescribed = prescribed_dofs(*p);
    // assembly.rs:55-81
    std::vector<std::vector<u32>> mapped;
    mapped.reserve(p->mesh.tets.size());
    for (const auto &tet : p->mesh.tets) {
        std::vector<u32> kept;
        for (u32 node : tet)
            if (p->free[node] != FIXED) kept.push_back(p->free[node]);
        mapped.push_back(std::move(kept));
    }
    p->matrix = CsrMatrix::from_block_connectivity(p->free_count, 3, mapped);
    // assembly.rs:87-130: element order
    double ke[900];
    for (size_t t = 0; t < p->mesh.tets.size(); ++t) {
        const auto &tet = p->mesh.tets[t];
        Vec3 en[10];
        for (int a = 0; a < 10; ++a) en[a] = nodes[tet[a]];
        if (!element_stiffness(en, p->material, p->quadrature, ke)) throw std::runtime_error("element fault");
        for (int a = 0; a < 10; ++a) {
            u32 rf = p->free[tet[a]];
            if (rf == FIXED) continue;
            for (int b = 0; b < 10; ++b) {
                u32 cf = p->free[tet[b]];
                if (cf == FIXED) continue;
                for (int i = 0; i < 3; ++i)
                    for (int k = 0; k < 3; ++k)
                        p->matrix.add(3 * (size_t)rf + i, (u32)(3 * (size_t)cf + k), ke[(3 * a + i) * 30 + (3 * b + k)]);
            }
        }
    }
    // study.rs:499-502: the elimination once every element is in
    if (!p->prescribed.empty()) apply_prescribed(*p);
    p->assembly_seconds = now_seconds() - t0;
    return p;
}

// study.rs:218-237: every mesh node's displacement, fixed nodes zero, prescribed components exact
static std::vector<double> nodal_displacements(const Problem &p, const std::vector<double> &x) {
    const size_t nn = p.mesh.nodes.size();
    std::vector<double> disp(3 * nn, 0.0);
    for (size_t n = 0; n < nn; ++n)
        if (p.free[n] != FIXED)
            for (int a = 0; a < 3; ++a) disp[3 * n + a] = x[3 * (size_t)p.free[n] + a];
    for (const auto &kv : p.prescribed_nodes) disp[3 * (size_t)kv.first.first + kv.first.second] = kv.second;
    return disp;
}

// study.rs:529-549: the rigid modes as near null space, zero on prescribed components
static std::vector<double> setup_modes(const Problem &p) {
    auto modes = free_rigid_body_modes(p.mesh.nodes, p.free, p.free_count);
    for (const auto &pr : p.prescribed) std::fill(&modes[pr.first * 6], &modes[pr.first * 6] + 6, 0.0);
    return modes;
}

// ---------------------------------------------------------------------------------------------
// PCG: crates/solidfem-solve/src/pcg.rs:113-301
// ---------------------------------------------------------------------------------------------

static double dot(const std::vector<double> &a, const std::vector<double> &b) {
    double sum = 0.0;
    for (size_t i = 0; i < a.size(); ++i) sum += a[i] * b[i];
    return sum;
}
static double norm(const std::vector<double> &a) { return std::sqrt(dot(a, a)); }
static void add_scaled(std::vector<double> &t, double scale, const std::vector<double> &d) {
    for (size_t i = 0; i < t.size(); ++i) t[i] += scale * d[i];
}

enum class Status { Active, Converged, Cancelled, NotPositiveDefinite, NotConverged };

template <typename Apply>
struct Pcg {
    const CsrMatrix &matrix;
    Apply precondition;
    // Preconditioner::system_multiply (preconditioner.rs:10-20): the matrix's own product unless
    // the preconditioner supplies one (the multigrid's, on its pool: cycle.rs:447-457)
    std::function<void(const double *, double *)> system_multiply;
    std::vector<double> rhs, x, r, z, p, ap;
    double rz = 0, rhs_norm = 0, relative_residual = 0, tolerance;
    u64 iterations = 0, max_iterations;
    bool floor_reached = false;
    Status status = Status::Active;

    // pcg.rs:135-164
    Pcg(const CsrMatrix &m, Apply apply, std::vector<double> b, double tol, u64 max_it)
        : matrix(m), precondition(apply), rhs(std::move(b)), tolerance(tol) {
        system_multiply = [this](const double *x, double *y) { matrix.multiply(x, y); };
        size_t n = matrix.size;
        rhs_norm = norm(rhs);
        max_iterations = max_it ? max_it : 10 * (u64)n + 1000;
        r = rhs;
        x.assign(n, 0.0);
        z.assign(n, 0.0);
        p.assign(n, 0.0);
        ap.assign(n, 0.0);
        if (rhs_norm == 0.0) {
            status = Status::Converged;
        } else {
            reset_direction();
            relative_residual = 1.0;
        }
    }
    void reset_direction() {
        precondition(r.data(), z.data());
        p = z;
        rz = dot(r, z);
    }
    // pcg.rs:168-189 (no cancel here: the oracle never cancels)
    Status step(u64 max_steps) {
        if (status != Status::Active) return status;
        for (u64 s = 0; s < max_steps; ++s) {
            iterate();
            if (status != Status::Active) return status;
        }
        return status;
    }
    // pcg.rs:216-275
    void iterate() {
        system_multiply(p.data(), ap.data());
        double curvature = dot(p, ap);
        if (std::isnan(curvature) || curvature <= 0.0 || !std::isfinite(rz)) {
            status = Status::NotPositiveDefinite;
            return;
        }
        double alpha = rz / curvature;
        add_scaled(x, alpha, p);
        add_scaled(r, -alpha, ap);
        iterations += 1;
        relative_residual = norm(r) / rhs_norm;
        if (relative_residual <= tolerance) {
            std::vector<double> residual(rhs.size(), 0.0);
            system_multiply(x.data(), residual.data());
            for (size_t i = 0; i < residual.size(); ++i) residual[i] = rhs[i] - residual[i];
            relative_residual = norm(residual) / rhs_norm;
            double floor = matrix.multiply_rounding_bound(x.data()) / rhs_norm;
            floor_reached = relative_residual > tolerance && relative_residual <= floor;
            if (relative_residual <= tolerance || floor_reached) {
                status = Status::Converged;
                return;
            }
            r = std::move(residual);
            reset_direction();
        } else {
            precondition(r.data(), z.data());
            double rz_next = dot(r, z);
            double beta = rz_next / rz;
            rz = rz_next;
            for (size_t i = 0; i < p.size(); ++i) p[i] = z[i] + beta * p[i];
        }
        if (iterations >= max_iterations) status = Status::NotConverged;
    }
};

static uint64_t fnv1a(const double *values, int64_t count) {
    uint64_t hash = 0xcbf29ce484222325ull;
    for (int64_t i = 0; i < count; ++i) {
        uint64_t bits;
        std::memcpy(&bits, &values[i], 8);
        for (int b = 0; b < 8; ++b) {
            hash ^= (bits >> (8 * b)) & 0xff;
            hash *= 0x00000100000001b3ull;
        }
    }
    return hash;
}

static MultigridOptions to_options(const ref_mg_options *o) {
    MultigridOptions m;
    m.neighbour_reach = o->neighbour_reach;
    m.coarse_dofs = (size_t)o->coarse_dofs;
    m.max_levels = (size_t)o->max_levels;
    m.smoother_degree = (size_t)o->smoother_degree;
    m.smoother_range = o->smoother_range;
    m.precision = (Precision)o->precision;
    m.threads = o->threads > 0 ? (size_t)o->threads : default_threads();
    return m;
}

}  // namespace refcpu

using namespace refcpu;

struct ref_problem {
    std::unique_ptr<Problem> p;
};
struct ref_multigrid {
    std::unique_ptr<Multigrid> m;
    double setup_seconds = 0;
};

extern "C" {

void ref_default_workload(ref_workload *w) {
    std::memset(w, 0, sizeof(*w));
    w->size[0] = w->size[1] = w->size[2] = 1.0;
    w->divisions[0] = w->divisions[1] = w->divisions[2] = 1;
    w->load_axis = 0;
    w->load_value = 1.0;
    w->traction_kind = REF_TRACTION_UNIFORM;
    w->material_kind = REF_MATERIAL_UNIFORM;
    w->youngs = 200000.0;
    w->poisson = 0.3;
    w->quadrature_points = 4;
}

void ref_default_mg_options(ref_mg_options *o) {
    o->neighbour_reach = 3.0;
    o->coarse_dofs = 1500;
    o->max_levels = 12;
    o->smoother_degree = 3;
    o->smoother_range = 30.0;
    o->precision = 1;
    o->threads = 0;
}

ref_problem *ref_problem_create(const ref_workload *w) {
    try {
        auto *rp = new ref_problem;
        rp->p = build_problem(*w);
        return rp;
    } catch (...) {
        return nullptr;
    }
}
ref_problem *ref_problem_create_prescribed(const ref_workload *w, const uint32_t *nodes, const uint32_t *axes, const double *values,
                                           int64_t count) {
    try {
        auto rp = std::make_unique<ref_problem>();
        rp->p = build_problem(*w, nodes, axes, values, count);
        return rp.release();
    } catch (...) {
        return nullptr;
    }
}
void ref_problem_destroy(ref_problem *p) { delete p; }

int64_t ref_problem_node_count(const ref_problem *p) { return (int64_t)p->p->mesh.nodes.size(); }
int64_t ref_problem_vertex_count(const ref_problem *p) { return (int64_t)p->p->mesh.vertex_count; }
int64_t ref_problem_tet_count(const ref_problem *p) { return (int64_t)p->p->mesh.tets.size(); }
int64_t ref_problem_free_dofs(const ref_problem *p) { return (int64_t)(3 * p->p->free_count); }
int64_t ref_problem_nonzeros(const ref_problem *p) { return (int64_t)p->p->matrix.columns.size(); }
int64_t ref_problem_boundary_face_count(const ref_problem *p) { return (int64_t)p->p->boundary.size(); }
double ref_problem_assembly_seconds(const ref_problem *p) { return p->p->assembly_seconds; }

void ref_problem_nodes(const ref_problem *p, double *out) {
    for (size_t n = 0; n < p->p->mesh.nodes.size(); ++n)
        for (int a = 0; a < 3; ++a) out[3 * n + a] = p->p->mesh.nodes[n][a];
}
void ref_problem_tets(const ref_problem *p, uint32_t *out) {
    for (size_t t = 0; t < p->p->mesh.tets.size(); ++t)
        for (int a = 0; a < 10; ++a) out[10 * t + a] = p->p->mesh.tets[t][a];
}
void ref_problem_free_map(const ref_problem *p, uint32_t *out) {
    std::copy(p->p->free.begin(), p->p->free.end(), out);
}
void ref_problem_boundary_faces(const ref_problem *p, uint32_t *out) {
    for (size_t f = 0; f < p->p->boundary.size(); ++f) {
        const auto &b = p->p->boundary[f];
        out[8 * f] = b.tet;
        out[8 * f + 1] = b.face;
        for (int s = 0; s < 6; ++s) out[8 * f + 2 + s] = b.nodes[s];
    }
}
int64_t ref_problem_loaded_face_count(const ref_problem *p) { return (int64_t)p->p->tractions.size(); }
void ref_problem_loaded_faces(const ref_problem *p, uint32_t *out, double *tractions) {
    for (size_t f = 0; f < p->p->tractions.size(); ++f) {
        out[2 * f] = p->p->tractions[f].face.tet;
        out[2 * f + 1] = p->p->tractions[f].face.face;
        for (int a = 0; a < 3; ++a) tractions[3 * f + a] = p->p->tractions[f].t[a];
    }
}
void ref_problem_csr(const ref_problem *p, int64_t *row_starts, uint32_t *columns, double *values) {
    const auto &m = p->p->matrix;
    for (size_t i = 0; i < m.row_starts.size(); ++i) row_starts[i] = (int64_t)m.row_starts[i];
    std::copy(m.columns.begin(), m.columns.end(), columns);
    std::copy(m.values.begin(), m.values.end(), values);
}
void ref_problem_rhs(const ref_problem *p, double *out) { std::copy(p->p->rhs.begin(), p->p->rhs.end(), out); }
void ref_problem_rigid_modes(const ref_problem *p, double *out) {
    auto modes = setup_modes(*p->p);
    std::copy(modes.begin(), modes.end(), out);
}
// Stress recovery of nodal displacements `u` (3 per node), results.rs:307-576: per (element,
// point) 10 values (position, stress, von Mises) and per node 7 (stress, von Mises). Returns 0,
// or -1 with the failing element in *fault_tet.
int ref_recover(const ref_problem *p, const double *u, double *point_values, double *nodal_values, int64_t *fault_tet) {
    const Problem &pr = *p->p;
    std::vector<double> disp(u, u + 3 * pr.mesh.nodes.size());
    RecoveredField f = recover(pr.mesh, pr.material, pr.quadrature, disp);
    if (f.fault_tet >= 0) {
        if (fault_tet) *fault_tet = f.fault_tet;
        return -1;
    }
    const size_t points = f.point_von_mises.size();
    if (point_values)
        for (size_t k = 0; k < points; ++k) {
            for (int a = 0; a < 3; ++a) point_values[10 * k + a] = f.point_position[3 * k + a];
            for (int c = 0; c < 6; ++c) point_values[10 * k + 3 + c] = f.point_stress[6 * k + c];
            point_values[10 * k + 9] = f.point_von_mises[k];
        }
    if (nodal_values)
        for (size_t n = 0; n < f.nodal_von_mises.size(); ++n) {
            for (int c = 0; c < 6; ++c) nodal_values[7 * n + c] = f.nodal_stress[6 * n + c];
            nodal_values[7 * n + 6] = f.nodal_von_mises[n];
        }
    return 0;
}
int ref_error_estimate(const ref_problem *p, const double *u, double *element_errors, double *sums, int64_t *payload) {
    const Problem &pr = *p->p;
    std::vector<double> disp(u, u + 3 * pr.mesh.nodes.size());
    const RecoveredField f = recover(pr.mesh, pr.material, pr.quadrature, disp);
    if (f.fault_tet >= 0) {
        payload[0] = f.fault_tet;
        payload[1] = f.fault_point;
        payload[2] = 0;
        return ESTIMATE_INVERTED;
    }
    const ErrorEstimate e = estimate_error(pr.mesh, pr.material, disp, f);
    std::copy(e.element_errors.begin(), e.element_errors.end(), element_errors);
    sums[0] = e.error_sum;
    sums[1] = e.energy_sum;
    sums[2] = e.relative_error;
    payload[0] = e.a;
    payload[1] = e.b;
    payload[2] = (int64_t)e.element_errors.size();
    return e.status;
}
int ref_problem_quadrature(const ref_problem *p) { return p->p->quadrature; }
// peak_with_holder (results.rs:184-226) over the candidates the masks keep (NULL: all). Returns 1
// with the peak, 0 when no candidate is inside.
int ref_peak(const double *positions, uint64_t nodes, const double *nodal_vm, uint64_t stride, const double *points, uint64_t count,
             const uint8_t *node_inside, const uint8_t *point_inside, int32_t *at_point, uint64_t *index, double *von_mises,
             double *position) {
    const Peak peak = peak_of(positions, (size_t)nodes, nodal_vm, (size_t)stride, points, (size_t)count, node_inside, point_inside);
    if (!peak.found) return 0;
    *at_point = peak.at_point ? 1 : 0;
    *index = peak.index;
    *von_mises = peak.von_mises;
    for (int a = 0; a < 3; ++a) position[a] = peak.position[a];
    return 1;
}
int ref_compare(const ref_problem *coarse, const double *coarse_nodal_vm, uint64_t coarse_stride, const double *coarse_points,
                uint64_t coarse_count, const uint8_t *coarse_node_inside, const uint8_t *coarse_point_inside, const ref_problem *fine,
                const double *fine_nodal_vm, uint64_t fine_stride, const double *fine_points, uint64_t fine_count,
                const uint8_t *fine_node_inside, const uint8_t *fine_point_inside, double *values, int64_t *integers) {
    const ComparedSide c{&coarse->p->mesh, coarse_nodal_vm, coarse->p->mesh.nodes.size(), (size_t)coarse_stride, coarse_points,
                         (size_t)coarse_count, coarse_node_inside, coarse_point_inside};
    const ComparedSide f{&fine->p->mesh, fine_nodal_vm, fine->p->mesh.nodes.size(), (size_t)fine_stride, fine_points,
                         (size_t)fine_count, fine_node_inside, fine_point_inside};
    const Comparison r = compare_meshes(c, f);
    const ComparedPeak *sides[2] = {&r.coarse, &r.fine};
    for (int w = 0; w < 2; ++w) {
        values[5 * w] = sides[w]->peak.von_mises;
        for (int a = 0; a < 3; ++a) values[5 * w + 1 + a] = sides[w]->peak.position[a];
        values[5 * w + 4] = sides[w]->volume;
        integers[3 * w] = (int64_t)sides[w]->elements;
        integers[3 * w + 1] = sides[w]->peak.at_point ? 1 : 0;
        integers[3 * w + 2] = (int64_t)sides[w]->peak.index;
    }
    values[10] = r.change;
    values[11] = r.distance;
    values[12] = r.value;
    integers[6] = r.a;
    integers[7] = r.b;
    return r.status;
}
// nodal_rotations (results.rs:262-305) of nodal displacements `u` (3 per node): 3 per node. Returns
// 0, or -1 with the failing element in *fault_tet.
int ref_nodal_rotations(const ref_problem *p, const double *u, double *out, int64_t *fault_tet) {
    const Problem &pr = *p->p;
    std::vector<double> disp(u, u + 3 * pr.mesh.nodes.size()), rotations;
    const int64_t fault = nodal_rotations(pr.mesh, disp, rotations);
    if (fault >= 0) {
        if (fault_tet) *fault_tet = fault;
        return -1;
    }
    std::copy(rotations.begin(), rotations.end(), out);
    return 0;
}

int ref_problem_element_stiffness(const ref_problem *p, int64_t tet, double *ke) {
    const auto &t = p->p->mesh.tets[(size_t)tet];
    Vec3 en[10];
    for (int a = 0; a < 10; ++a) en[a] = p->p->mesh.nodes[t[a]];
    return element_stiffness(en, p->p->material, p->p->quadrature, ke) ? 0 : -1;
}

ref_multigrid *ref_multigrid_create(const ref_problem *rp, const ref_mg_options *o, int32_t *status) {
    const Problem &p = *rp->p;
    double t0 = now_seconds();
    try {
        // study.rs:530-549
        auto modes = setup_modes(p);
        std::vector<Vec3> positions;
        positions.reserve(p.free_count);
        for (size_t n = 0; n < p.mesh.nodes.size(); ++n)
            if (p.free[n] != FIXED) positions.push_back(p.mesh.nodes[n]);
        MultigridOptions opt = to_options(o);
        auto levels = build_hierarchy(p.matrix, modes, positions, opt);
        auto *m = new ref_multigrid;
        m->m = std::make_unique<Multigrid>(std::move(levels), opt.smoother_degree, opt.precision, opt.threads);
        m->setup_seconds = now_seconds() - t0;
        if (status) *status = 0;
        return m;
    } catch (...) {
        if (status) *status = 3;
        return nullptr;
    }
}
void ref_multigrid_destroy(ref_multigrid *m) { delete m; }
int64_t ref_multigrid_levels(const ref_multigrid *m) { return (int64_t)m->m->levels64.size(); }
int64_t ref_multigrid_threads(const ref_multigrid *m) { return (int64_t)m->m->threads(); }
int64_t ref_default_threads(void) { return (int64_t)default_threads(); }
double ref_multigrid_setup_seconds(const ref_multigrid *m) { return m->setup_seconds; }
void ref_multigrid_level_info(const ref_multigrid *m, int64_t level, int64_t *info) {
    const auto &l = m->m->levels64[(size_t)level];
    info[0] = (int64_t)l.a.rows;
    info[1] = (int64_t)l.a.row_block;
    info[2] = (int64_t)l.a.columns.size();
    info[3] = l.has_prolongator ? (int64_t)l.prolongator.columns.size() : 0;
    info[4] = l.has_coarse ? 1 : 0;
}
static void copy_block(const BlockCsr<double> &b, int64_t *starts, uint32_t *columns, double *values) {
    for (size_t i = 0; i < b.starts.size(); ++i) starts[i] = (int64_t)b.starts[i];
    std::copy(b.columns.begin(), b.columns.end(), columns);
    std::copy(b.values.begin(), b.values.end(), values);
}
void ref_multigrid_level_matrix(const ref_multigrid *m, int64_t level, int64_t *starts, uint32_t *columns, double *values) {
    copy_block(m->m->levels64[(size_t)level].a, starts, columns, values);
}
void ref_multigrid_level_prolongator(const ref_multigrid *m, int64_t level, int64_t *starts, uint32_t *columns, double *values) {
    copy_block(m->m->levels64[(size_t)level].prolongator, starts, columns, values);
}
void ref_multigrid_level_inverse_diagonal(const ref_multigrid *m, int64_t level, double *out) {
    const auto &d = m->m->levels64[(size_t)level].inverse_diagonal;
    std::copy(d.begin(), d.end(), out);
}
void ref_multigrid_level_spectrum(const ref_multigrid *m, int64_t level, double *mm) {
    mm[0] = m->m->levels64[(size_t)level].lambda_min;
    mm[1] = m->m->levels64[(size_t)level].lambda_max;
}
void ref_multigrid_level_aggregates(const ref_multigrid *m, int64_t level, uint32_t *out) {
    const auto &a = m->m->levels64[(size_t)level].aggregates;
    std::copy(a.begin(), a.end(), out);
}
int ref_multigrid_coarse_factor(const ref_multigrid *m, double *out) {
    const auto &l = m->m->levels64.back();
    if (!l.has_coarse) return 0;
    std::copy(l.coarse.factor.begin(), l.coarse.factor.end(), out);
    return 1;
}
void ref_multigrid_apply(const ref_multigrid *m, const double *r, double *z) { m->m->apply(r, z); }
void ref_multigrid_smooth(const ref_multigrid *m, int64_t level, const double *b, double *x) {
    m->m->smooth_level((size_t)level, b, x);
}
void ref_multigrid_stats(const ref_multigrid *m, int64_t *levels, int64_t *coarse_dofs, double *oc, int64_t *memory) {
    size_t l, c, mem;
    double o;
    m->m->stats(l, c, o, mem);
    *levels = (int64_t)l;
    *coarse_dofs = (int64_t)c;
    *oc = o;
    *memory = (int64_t)mem;
}

uint64_t ref_fingerprint(const double *values, int64_t count) { return fnv1a(values, count); }

static CsrMatrix csr_view(int64_t n, const int64_t *rs, const uint32_t *cols, const double *vals) {
    CsrMatrix m;
    m.size = (size_t)n;
    m.row_starts.assign(rs, rs + n + 1);
    m.columns.assign(cols, cols + rs[n]);
    m.values.assign(vals, vals + rs[n]);
    return m;
}
void ref_csr_multiply(int64_t n, const int64_t *rs, const uint32_t *cols, const double *vals, const double *x, double *y) {
    csr_view(n, rs, cols, vals).multiply(x, y);
}
double ref_csr_rounding_bound(int64_t n, const int64_t *rs, const uint32_t *cols, const double *vals, const double *x) {
    return csr_view(n, rs, cols, vals).multiply_rounding_bound(x);
}

int ref_solve_with(const ref_problem *rp, const ref_multigrid *rm, double tolerance, uint64_t iteration_limit,
                   double *displacements, ref_result *result) {
    const Problem &p = *rp->p;
    const Multigrid &mg = *rm->m;
    auto apply = [&mg](const double *r, double *z) { mg.apply(r, z); };
    double t0 = now_seconds();
    Pcg<decltype(apply)> pcg(p.matrix, apply, p.rhs, tolerance, 0);
    pcg.system_multiply = [&](const double *x, double *y) { mg.system_multiply(p.matrix, x, y); };
    Status st = pcg.step(iteration_limit ? iteration_limit : ~(u64)0);
    double t1 = now_seconds();
    size_t nn = p.mesh.nodes.size();
    std::vector<double> disp = nodal_displacements(p, pcg.x);
    if (displacements) std::copy(disp.begin(), disp.end(), displacements);
    if (result) {
        std::memset(result, 0, sizeof(*result));
        result->status = st == Status::Converged ? 0 : st == Status::NotConverged ? 1 : st == Status::NotPositiveDefinite ? 2 : 5;
        result->iterations = pcg.iterations;
        result->relative_residual = pcg.relative_residual;
        result->at_rounding_floor = pcg.floor_reached ? 1 : 0;
        result->fingerprint = fnv1a(disp.data(), (int64_t)disp.size());
        result->total_dofs = (int64_t)(3 * nn);
        result->free_dofs = (int64_t)(3 * p.free_count);
        result->nonzeros = (int64_t)p.matrix.columns.size();
        size_t l, c, mem;
        double o;
        mg.stats(l, c, o, mem);
        result->levels = (int64_t)l;
        result->coarse_dofs = (int64_t)c;
        result->operator_complexity = o;
        result->memory_bytes = (int64_t)mem;
        result->assembly_seconds = p.assembly_seconds;
        result->setup_seconds = rm->setup_seconds;
        result->solve_seconds = t1 - t0;
    }
    return result ? result->status : 0;
}

int ref_solve_rhs(const ref_problem *rp, const ref_multigrid *rm, const double *rhs, double tolerance, double *x, ref_result *result) {
    // Pcg::new + step (pcg.rs:135-189) on the problem's matrix with another right-hand side
    const Problem &p = *rp->p;
    const Multigrid &mg = *rm->m;
    auto apply = [&mg](const double *r, double *z) { mg.apply(r, z); };
    Pcg<decltype(apply)> pcg(p.matrix, apply, std::vector<double>(rhs, rhs + p.matrix.size), tolerance, 0);
    pcg.system_multiply = [&](const double *v, double *y) { mg.system_multiply(p.matrix, v, y); };
    Status st = pcg.step(~(u64)0);
    std::copy(pcg.x.begin(), pcg.x.end(), x);
    if (result) {
        std::memset(result, 0, sizeof(*result));
        result->status = st == Status::Converged ? 0 : st == Status::NotConverged ? 1 : 2;
        result->iterations = pcg.iterations;
        result->relative_residual = pcg.relative_residual;
        result->at_rounding_floor = pcg.floor_reached ? 1 : 0;
    }
    return result ? result->status : 0;
}

int ref_solve(const ref_workload *w, const ref_mg_options *o, double tolerance, double *displacements, ref_result *result) {
    ref_problem *p = ref_problem_create(w);
    if (!p) {
        if (result) {
            std::memset(result, 0, sizeof(*result));
            result->status = 4;
        }
        return 4;
    }
    int32_t st = 0;
    ref_multigrid *m = ref_multigrid_create(p, o, &st);
    if (!m) {
        ref_problem_destroy(p);
        if (result) {
            std::memset(result, 0, sizeof(*result));
            result->status = 3;
        }
        return 3;
    }
    int r = ref_solve_with(p, m, tolerance, 0, displacements, result);
    ref_multigrid_destroy(m);
    ref_problem_destroy(p);
    return r;
}

int ref_pcg_jacobi(int64_t n, const int64_t *row_starts, const uint32_t *columns, const double *values,
                   const double *rhs, double tolerance, uint64_t slice, double *x, ref_result *result) {
    CsrMatrix m;
    m.size = (size_t)n;
    m.row_starts.assign(row_starts, row_starts + n + 1);
    size_t nnz = (size_t)row_starts[n];
    m.columns.assign(columns, columns + nnz);
    m.values.assign(values, values + nnz);
    // preconditioner.rs:53-66
    std::vector<double> inv(n);
    for (int64_t i = 0; i < n; ++i) {
        double d = m.get((size_t)i, (u32)i);
        if (std::isnan(d) || d <= 0.0 || !std::isfinite(d)) return 2;
        inv[i] = 1.0 / d;
    }
    auto apply = [&inv](const double *r, double *z) {
        for (size_t i = 0; i < inv.size(); ++i) z[i] = r[i] * inv[i];
    };
    Pcg<decltype(apply)> pcg(m, apply, std::vector<double>(rhs, rhs + n), tolerance, 0);
    Status st;
    do {
        st = pcg.step(slice ? slice : ~(u64)0);
    } while (st == Status::Active);
    std::copy(pcg.x.begin(), pcg.x.end(), x);
    if (result) {
        std::memset(result, 0, sizeof(*result));
        result->status = st == Status::Converged ? 0 : st == Status::NotConverged ? 1 : 2;
        result->iterations = pcg.iterations;
        result->relative_residual = pcg.relative_residual;
        result->at_rounding_floor = pcg.floor_reached ? 1 : 0;
        result->fingerprint = fnv1a(x, n);
    }
    return result ? result->status : 0;
}

}  // extern "C"
