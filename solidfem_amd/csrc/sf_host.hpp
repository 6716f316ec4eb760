// Host-side structures of the solve path: meshes, the model, the free numbering, the load, the
// rigid-body modes and the block sparsity pattern with its assembly map. These feed the device
// (assembly, multigrid cycle, conjugate gradient) and the host multigrid setup.
#pragma once

#include <functional>
#include <map>
#include <set>

#include "sf_common.hpp"

namespace sf {

struct Vec3 {
    double v[3];
    double &operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
};

// ---- meshes ------------------------------------------------------------------------------

// Four-node tetrahedra, validated as TetMesh::new does (crates/solidfem-mesh/src/tet.rs:59-109).
struct TetMesh {
    std::vector<Vec3> positions;
    std::vector<std::array<u32, 4>> tets;
};
void validate_tet_mesh(const TetMesh &mesh);  // throws Error(SF_ERR_MESH_*)
// A mesh from its text form: `vertex-count tet-count`, x y z per vertex, a b c d per tetrahedron,
// validated; malformed text throws Error(SF_ERR_INVALID_ARGUMENT, a = line).
TetMesh read_tet_mesh_text(const char *text, size_t length);
TetMesh make_box_mesh(const double origin[3], const double size[3], const int32_t divisions[3]);
// The tube of a workload with wall_cells > 0 (demo/wasm/src/part.rs:48-185).
TetMesh make_tube_mesh(const sf_workload &w);

struct Face {
    u32 tet;
    u32 face;
    u32 nodes[6];
    bool operator<(const Face &o) const {
        if (tet != o.tet) return tet < o.tet;
        if (face != o.face) return face < o.face;
        for (int i = 0; i < 6; ++i)
            if (nodes[i] != o.nodes[i]) return nodes[i] < o.nodes[i];
        return false;
    }
    bool operator==(const Face &o) const {
        if (tet != o.tet || face != o.face) return false;
        for (int i = 0; i < 6; ++i)
            if (nodes[i] != o.nodes[i]) return false;
        return true;
    }
};

// Ten-node tetrahedra: corners then edge midpoints numbered by sorted corner pair
// (crates/solidfem-mesh/src/quadratic.rs:24-71).
struct QuadMesh {
    std::vector<Vec3> nodes;
    std::vector<std::array<u32, 10>> tets;
    std::vector<std::array<u32, 4>> corners;
    size_t vertex_count = 0;
    std::vector<Face> boundary;  // faces met once, in quadratic.rs:89-124 order
    std::vector<std::array<u32, 2>> shared_faces;  // tetrahedra meeting at a corner triangle
    const std::vector<Face> &boundary_faces() const { return boundary; }
};
QuadMesh upgrade_to_quadratic(const TetMesh &mesh);
// The elements grouped into bodies joined through shared faces (quadratic.rs:129-166).
std::vector<std::vector<u32>> face_connected_bodies(const QuadMesh &mesh);

// Local node indices of each face (crates/solidfem-mesh/src/boundary.rs:8-13).
extern const int FACE_LOCAL[4][6];
extern const int EDGE_PAIRS[6][2];

// ---- model -------------------------------------------------------------------------------

struct MaterialField {
    // Uniform (E, nu), a shell of depth `shell_depth` below the two long-face pairs with a core, or
    // a caller's field evaluated at every point of both rules (SF_MATERIAL_POINTS).
    int kind = SF_MATERIAL_UNIFORM;
    double e = 1, nu = 0, core_e = 1, core_nu = 0, shell_depth = 0;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    int transverse[2] = {1, 2};
    // SF_MATERIAL_POINTS: (E, nu) per (element, point) of the 4- and the 14-point rule
    std::shared_ptr<const std::vector<double>> points4, points14;
    const std::vector<double> &points(int nq) const { return nq == 14 ? *points14 : *points4; }
};

struct Model {
    QuadMesh mesh;
    MaterialField field;
    int quadrature = 4;
    std::set<u32> fixed;
    std::vector<std::pair<u32, std::array<double, 3>>> nodal_forces;
    std::vector<std::pair<Face, std::array<double, 3>>> tractions;
    std::map<std::pair<u32, u32>, double> prescribed;
};

Model model_from_workload(const sf_workload &w);

using Traction = std::pair<Face, std::array<double, 3>>;
// The tractions a workload puts on the boundary faces of its load plane; `boundary` lists the
// mesh's boundary faces in QuadMesh::boundary_faces order (random tractions are drawn by it).
std::vector<Traction> workload_tractions(const QuadMesh &mesh, const std::vector<Face> &boundary, const sf_workload &w);

// The validated, numbered problem of StudyBuilder::new (crates/solidfem-elastic/src/study.rs:419-474).
struct FreeProblem {
    std::vector<u32> free;  // node -> free index or FIXED
    size_t free_count = 0;
    std::vector<double> rhs;  // 3 per free node
    std::vector<std::pair<size_t, double>> prescribed;  // (free dof, value), sorted
};
FreeProblem number_and_load(const Model &model);  // throws Error(SF_ERR_MODEL_*)
// The nodal load of a model's forces and tractions (study.rs:696-735), 3 per mesh node; the
// mesh's boundary faces, sorted, may be given to save finding them again.
std::vector<double> model_load(const Model &model, const std::vector<Face> *sorted_boundary = nullptr);
// The right-hand side over the free unknowns of tractions alone: number_and_load's load for a
// model with no nodal forces and no prescribed displacements, the same bits.
std::vector<double> traction_rhs(const QuadMesh &mesh, const std::vector<Traction> &tractions, const std::vector<u32> &free,
                                 size_t free_count);

// Six rigid-body modes per free unknown (crates/solidfem-elastic/src/rigid_modes.rs:36-83).
std::vector<double> rigid_body_modes(const QuadMesh &mesh, const std::vector<u32> &free, size_t free_count);

// ---- block pattern and assembly map ----------------------------------------------------

// A block sparse matrix on the host: rows of `rb` x `cb` blocks, sorted block columns,
// row-major block values.
template <typename T>
struct HostBsr {
    size_t rows = 0, cols = 0, rb = 0, cb = 0;
    std::vector<int64_t> starts{0};
    std::vector<u32> columns;
    std::vector<T> values;
    size_t blocks() const { return columns.size(); }
    size_t block_len() const { return rb * cb; }
    size_t dofs() const { return rows * rb; }
    const T *block(size_t at) const { return &values[at * rb * cb]; }
};

// The free-node block pattern (the pattern of CsrMatrix::from_block_connectivity over the
// elements' free nodes, crates/solidfem-elastic/src/assembly.rs:55-81), built row by row from
// the node-to-element adjacency, and the assembly map: for every stored block, its element
// contributions in ascending element order, packed as (element << 7) | (10 a + b).
struct AssemblyPattern {
    HostBsr<double> pattern;               // values empty
    HostArray<int64_t> contribution_starts;  // blocks + 1
    HostArray<u32> contributions;
};
AssemblyPattern build_pattern(const QuadMesh &mesh, const std::vector<u32> &free, size_t free_count);

// study.rs:567-593: each prescribed row's off-diagonal entries, read before that row's own updates,
// move `entry * value` off the right-hand side of every unconstrained row and are cancelled in
// both the row and its mirror column; then each prescribed row's right-hand side becomes its
// diagonal times the value. The system stays symmetric positive definite with the prescribed
// unknowns decoupled.
void apply_prescribed(HostBsr<double> &a, std::vector<double> &rhs, const std::vector<std::pair<size_t, double>> &prescribed);
// The right-hand side part of that elimination, recorded from the matrix before it, so that another
// load on the same system (the same mesh, constraints and prescribed values) gets the same bits:
// for each prescribed unknown in order, rhs[c] -= K[dof][c] * value over its row's unconstrained
// columns c, then every prescribed rhs[dof] = K[dof][dof] * value.
struct PrescribedRhs {
    std::vector<std::pair<size_t, double>> prescribed;  // (free dof, value), sorted
    std::vector<int64_t> starts;                        // entries of prescribed unknown k
    std::vector<std::pair<size_t, double>> entries;     // (column, K[dof][column]) before elimination
    std::vector<double> diagonal;                        // K[dof][dof]
};
PrescribedRhs prescribed_rhs_program(const HostBsr<double> &a, const std::vector<std::pair<size_t, double>> &prescribed);
void apply_prescribed_rhs(const PrescribedRhs &program, std::vector<double> &rhs);

// peak_with_holder (crates/solidfem-elastic/src/results.rs:184-226): the largest von Mises over
// the nodes, then the integration points, whose positions `inside` accepts. A later candidate
// replaces the peak only when strictly larger, so the first of equal values holds it; the first
// NaN met becomes the peak and stays it. Nodal values are read at nodal_vm[stride * node]; a point
// is 10 values (position 3, stress 6, von Mises).
struct PeakStress {
    bool found = false;
    bool at_point = false;  // held by an integration point, else by a mesh node
    size_t index = 0;       // of that node or point
    double von_mises = 0.0;
    double position[3] = {0.0, 0.0, 0.0};
};
PeakStress peak_stress(const double *positions, size_t nodes, const double *nodal_vm, size_t stride, const double *point_values,
                       size_t points, const std::function<bool(const double *)> &inside);

// One finished solve of a mesh comparison (crates/solidfem-elastic/src/comparison.rs): its mesh,
// its nodal von Mises (`nodes` values, `stride` apart) and its integration-point values (10 per
// point, the same number per element, in element order).
struct ComparedSolve {
    const QuadMesh *mesh = nullptr;
    const double *nodal_vm = nullptr;
    size_t nodes = 0, stride = 1;
    const double *points = nullptr;
    size_t point_count = 0;
};
struct ComparedPeak {
    PeakStress peak;
    size_t elements = 0;             // elements with an integration point inside the region
    double peak_element_volume = 0;  // the volume of the element(s) holding the peak
};
struct MeshComparison {
    ComparedPeak coarse, fine;
    double peak_change = 0, peak_distance = 0;
};
// MeshComparison::between (comparison.rs:101-141); throws the refusals
// (SF_ERR_MODEL_SOLUTION_MESH_MISMATCH, _NOTHING_INSIDE_REGION, _NON_FINITE_PEAK, _ZERO_PEAK,
// _FINE_MESH_NOT_FINER).
MeshComparison compare_meshes(const ComparedSolve &coarse, const ComparedSolve &fine,
                              const std::function<bool(const double *)> &inside);

}  // namespace sf
