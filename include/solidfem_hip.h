/*
 * solidfem_hip.h -- the C ABI of the MI355X-native linear-static solve path.
 *
 * The entry points replace the Rust API of solidfem's solve path one for one; each
 * declaration cites the reference item it stands in for (paths relative to the reference
 * repository). Conventions:
 *   - Opaque handles with explicit create/destroy. Host arrays are borrowed for the call and
 *     copied; the library owns all device memory.
 *   - Every call returns an sf_status. SF_OK is 0. On failure sf_last_error() holds a message
 *     and sf_last_error_payload() the named fields of the reference's error variant.
 *   - A handle is used by one thread at a time (the reference's Multigrid is !Sync:
 *     crates/solidfem-solve/src/multigrid/cycle.rs:111-114). Different handles may run
 *     concurrently on different streams.
 *   - The cancel callback runs on the calling thread, before every iteration of a solve
 *     (crates/solidfem-solve/src/pcg.rs:179-183) and between row chunks of a setup
 *     (crates/solidfem-solve/src/multigrid/setup.rs:191-192). Nothing throws across the ABI.
 *   - Results never depend on how a resumable step is sliced
 *     (crates/solidfem-solve/src/lib.rs:15-19).
 */
#ifndef SOLIDFEM_HIP_H
#define SOLIDFEM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t sf_status;

enum {
    SF_OK = 0,
    SF_ERR_INVALID_ARGUMENT = 1,
    SF_ERR_HIP = 2,
    SF_ERR_OUT_OF_MEMORY = 3,
    SF_ERR_INTERNAL = 4,
    SF_ERR_GRAPH_UNDER_TRACER = 5,       /* sf_multigrid_set_graph under rocprofv3 */
    /* PcgError, crates/solidfem-solve/src/pcg.rs:59-77 */
    SF_ERR_PCG_NOT_POSITIVE_DEFINITE = 10, /* payload: a=iteration, value=curvature */
    SF_ERR_PCG_NOT_CONVERGED = 11,         /* payload: a=iterations, value=relative residual */
    /* MultigridError, crates/solidfem-solve/src/multigrid/mod.rs:197-246 */
    SF_ERR_MG_BLOCK_STRUCTURE = 20,        /* a=row */
    SF_ERR_MG_NEAR_NULL_LENGTH = 21,       /* a=expected, b=found */
    SF_ERR_MG_NOT_POSITIVE_DEFINITE = 22,  /* a=level, b=block, value=pivot */
    SF_ERR_MG_CANCELLED = 23,
    SF_ERR_MG_POSITION_LENGTH = 24,        /* a=expected, b=found */
    SF_ERR_MG_NOT_FINISHED = 25,
    SF_ERR_MG_COARSEST_NOT_FACTORED = 26,  /* a=values */
    SF_ERR_MG_COARSE_LENGTH = 27,          /* a=expected, b=found */
    /* JacobiError, crates/solidfem-solve/src/preconditioner.rs:22-31 */
    SF_ERR_JACOBI_DIAGONAL = 28,           /* a=row, value=diagonal */
    SF_ERR_PRECONDITIONER = 29,            /* a caller's preconditioner failed: a=its return value */
    /* MeshError, crates/solidfem-mesh/src/error.rs:7-27 */
    SF_ERR_MESH_EMPTY = 30,
    SF_ERR_MESH_NON_FINITE_POSITION = 31,  /* a=node */
    SF_ERR_MESH_NODE_OUT_OF_RANGE = 32,    /* a=tet, b=node */
    SF_ERR_MESH_UNUSED_NODE = 33,          /* a=node */
    SF_ERR_MESH_DEGENERATE_TET = 34,       /* a=tet, value=volume */
    SF_ERR_MESH_FLAT_TET = 35,             /* a=tet, value=relative volume */
    SF_ERR_MESH_NON_MANIFOLD_FACE = 36,
    SF_ERR_MESH_INVALID_BOX = 37,
    /* ModelError, crates/solidfem-elastic/src/error.rs:23-95 (the solve path's variants) */
    SF_ERR_MODEL_NO_FIXED_NODES = 40,
    SF_ERR_MODEL_NODE_OUT_OF_RANGE = 41,   /* a=node */
    SF_ERR_MODEL_NON_FINITE_LOAD = 42,
    SF_ERR_MODEL_FACE_NOT_ON_BOUNDARY = 43,/* a=tet, b=face */
    SF_ERR_MODEL_NO_EFFECTIVE_LOAD = 44,
    SF_ERR_MODEL_INVERTED_JACOBIAN = 45,   /* a=tet, b=point, value=determinant */
    SF_ERR_MODEL_INVALID_MATERIAL = 46,    /* a=tet, b=point */
    SF_ERR_MODEL_CANCELLED = 47,
    SF_ERR_MODEL_NOT_FINISHED = 48,
    SF_ERR_MODEL_AXIS_OUT_OF_RANGE = 49,   /* a=axis */
    SF_ERR_MODEL_PRESCRIBED_ON_FIXED = 50, /* a=node */
    SF_ERR_MODEL_BODY_NOT_HELD = 51,       /* a=first tet, b=held nodes, value=SF_FREEDOM_* left */
    SF_ERR_MODEL_DOFS_DO_NOT_HOLD = 52,    /* a=first tet, b=rigid motions held (rank < 6) */
    SF_ERR_MODEL_MIXED_MATERIAL_PATCH = 53,/* a=vertex */
    SF_ERR_MODEL_COPLANAR_PATCH = 54,      /* a=vertex */
    SF_ERR_MODEL_ZERO_STRAIN_ENERGY = 55,
    SF_ERR_MODEL_NON_FINITE_ENERGY = 56,
    SF_ERR_MODEL_RECOVERY_MESH_MISMATCH = 57,
    SF_ERR_MODEL_SOLUTION_MESH_MISMATCH = 58, /* a=SF_COMPARED_* */
    SF_ERR_MODEL_NOTHING_INSIDE_REGION = 59,  /* a=SF_COMPARED_* */
    /* serving */
    SF_ERR_SERVE_OVER_BUDGET = 60,         /* a=bytes needed, b=budget */
    SF_ERR_SERVE_UNKNOWN_REQUEST = 61,
    /* ModelError's mesh-comparison refusals, continued (crates/solidfem-elastic/src/error.rs:95-119) */
    SF_ERR_MODEL_NON_FINITE_PEAK = 70,        /* a=SF_COMPARED_*, value=the peak */
    SF_ERR_MODEL_ZERO_PEAK = 71,              /* a=SF_COMPARED_* */
    SF_ERR_MODEL_FINE_MESH_NOT_FINER = 72,    /* a=coarse elements, b=fine elements in the region */
    SF_ERR_MODEL_INVALID_YIELD_STRENGTH = 73  /* value=the yield strength */
};

/* Which solve of a mesh comparison a refusal names (ComparedMesh, error.rs). */
enum { SF_COMPARED_COARSE = 0, SF_COMPARED_FINE = 1 };

/* The rigid motion an insufficiently held body keeps (Freedom,
 * crates/solidfem-elastic/src/constraints.rs:14-23). */
enum { SF_FREEDOM_FREE = 0, SF_FREEDOM_ROTATION_ABOUT_POINT = 1, SF_FREEDOM_ROTATION_ABOUT_LINE = 2 };

/* Precision of the finished hierarchy, crates/solidfem-solve/src/multigrid/mod.rs:126-141 */
enum { SF_PRECISION_DOUBLE = 0, SF_PRECISION_SINGLE_STORAGE = 1, SF_PRECISION_SINGLE = 2 };

/* Order of the conjugate gradient's dot products. REFERENCE sums in index order from zero,
 * exactly crates/solidfem-solve/src/pcg.rs:285-301, so a solve reproduces the reference's bits
 * (a serial reduction: for parity runs on small systems). TREE sums in a fixed wavefront tree
 * and a fixed order of partials: the same bits on every run, at full device speed. */
enum { SF_DOT_TREE = 0, SF_DOT_REFERENCE = 1 };

/* Step outcome, crates/solidfem-solve/src/pcg.rs:47-57 */
enum { SF_RUNNING = 0, SF_CONVERGED = 1, SF_CANCELLED = 2 };

/* A column of a batched solve that failed: the error is in its sf_column_progress. */
enum { SF_FAILED = 3 };

/* Returns non-zero to cancel. */
typedef int32_t (*sf_cancel_fn)(void *user);

/* Progress, crates/solidfem-solve/src/pcg.rs:35-45, plus the outcome of the call. */
typedef struct sf_progress {
    uint64_t iteration;
    double relative_residual;
    int32_t at_rounding_floor;
    int32_t outcome;
} sf_progress;

/* One column of a batched solve: its Progress (pcg.rs:35-45) and outcome; for SF_FAILED, the
 * PcgError variant as a status (SF_ERR_PCG_NOT_POSITIVE_DEFINITE or SF_ERR_PCG_NOT_CONVERGED)
 * with its iteration and value (curvature or relative residual). */
typedef struct sf_column_progress {
    uint64_t iteration;
    double relative_residual;
    int32_t at_rounding_floor;
    int32_t outcome;
    sf_status error;
    int32_t reserved;
    uint64_t error_iteration;
    double error_value;
} sf_column_progress;

/* PcgOptions, crates/solidfem-solve/src/pcg.rs:14-33 (max_iterations 0 = 10 n + 1000) */
typedef struct sf_pcg_options {
    double tolerance;
    uint64_t max_iterations;
    int32_t dot_order;
    int32_t reserved;
} sf_pcg_options;

/* MultigridOptions, crates/solidfem-solve/src/multigrid/mod.rs:143-182 */
typedef struct sf_multigrid_options {
    double neighbour_reach;
    uint64_t coarse_dofs;
    uint64_t max_levels;
    uint64_t smoother_degree;
    double smoother_range;
    int32_t precision;
    int32_t threads; /* host threads for setup; 0 = all */
    int32_t coarse_solve; /* SF_COARSE_REFERENCE or SF_COARSE_INVERSE */
    int32_t setup_device; /* SF_SETUP_AUTO, SF_SETUP_HOST or SF_SETUP_DEVICE */
} sf_multigrid_options;

/* Where a level's numeric setup phases run: the power iteration, the smoothed prolongator, A P
 * and P^T (A P) (setup.rs:390-449, :527-647). AUTO runs them on the device for levels of at
 * least 4096 block rows when the setup has a device context, HOST never, DEVICE for every
 * level (given a context). Aggregation, the tentative prolongator and the block inverses stay
 * on the host. The hierarchy is the same to the bit wherever the phases run. No counterpart in
 * the reference, whose setup runs on its host thread pool. */
enum { SF_SETUP_AUTO = 0, SF_SETUP_HOST = 1, SF_SETUP_DEVICE = 2 };

/* The coarsest level's solve. REFERENCE runs forward then back substitution with the Cholesky
 * factor in the reference's order (crates/solidfem-solve/src/multigrid/block.rs:357-373), so a
 * cycle repeats the reference's bits; its back substitution is a serial chain of n^2/2
 * subtractions. INVERSE applies the f64 inverse L^-T L^-1, formed once at setup, as a dense
 * product summed in a fixed order: the same bits on every run, differing from REFERENCE by
 * rounding only, in a few microseconds. */
enum { SF_COARSE_REFERENCE = 0, SF_COARSE_INVERSE = 1 };

/* MultigridStats, crates/solidfem-solve/src/multigrid/mod.rs:184-195 */
typedef struct sf_multigrid_stats {
    uint64_t levels;
    uint64_t coarse_dofs;
    double operator_complexity;
    uint64_t memory_bytes;
} sf_multigrid_stats;

/* Report, crates/solidfem-elastic/src/study.rs:150-176 */
typedef struct sf_report {
    uint64_t total_dofs;
    uint64_t free_dofs;
    uint64_t nonzeros;
    int32_t at_rounding_floor;
    int32_t has_multigrid;
    uint64_t iterations;
    double relative_residual;
    double assembly_seconds;
    double preconditioner_seconds;
    sf_multigrid_stats multigrid;
    double solve_seconds;
    double recovery_seconds; /* the study's recoveries (Report::recovery_seconds, study.rs:175) */
} sf_report;

/* BuildProgress, crates/solidfem-elastic/src/study.rs:377-387 */
typedef struct sf_build_progress {
    uint64_t elements_done;
    uint64_t elements_total;
    int32_t setup_finished;
    int32_t reserved;
} sf_build_progress;

/* The structured box problems the benchmark configurations are built from: a box mesh
 * (crates/solidfem-mesh/src/box_mesh.rs:22-80) with optional seeded interior jitter, upgraded to
 * Tet10, clamped on a plane, loaded by tractions on the boundary faces of a plane. clamp_axis -1
 * clamps no plane and load_axis -1 loads none (the model may then be held and loaded by
 * sf_model_fix_node, sf_model_prescribe_displacement and sf_model_add_traction); any other axis
 * outside 0..2 is refused with SF_ERR_MODEL_AXIS_OUT_OF_RANGE. */
enum { SF_TRACTION_UNIFORM = 0, SF_TRACTION_RANDOM = 1, SF_TRACTION_TORSION = 2 };
enum { SF_MATERIAL_UNIFORM = 0, SF_MATERIAL_SHELL_CORE = 1, SF_MATERIAL_POINTS = 2 };
typedef struct sf_workload {
    double origin[3];
    double size[3];
    int32_t divisions[3];
    double jitter_fraction;
    uint64_t jitter_seed;
    int32_t clamp_axis;
    double clamp_value;
    int32_t load_axis;
    double load_value;
    int32_t traction_kind;
    double traction[3];
    double traction_scale;
    uint64_t traction_seed;
    int32_t load_case;
    double axis_u, axis_v;
    int32_t material_kind;
    double youngs, poisson;
    double core_youngs, core_poisson;
    double shell_depth;
    int32_t quadrature_points;
    /* A tube when wall_cells > 0 (demo/wasm/src/part.rs:48-185): along the longest axis, the
     * two axes across it get wall_cells cells through each wall of wall_thickness and their
     * remaining divisions evenly over the hollow, whose cells are removed. Not jittered. */
    int32_t wall_cells;
    double wall_thickness;
} sf_workload;

typedef struct sf_context sf_context;
typedef struct sf_tet_mesh sf_tet_mesh;
typedef struct sf_quad_mesh sf_quad_mesh;
typedef struct sf_model sf_model;
typedef struct sf_study_builder sf_study_builder;
typedef struct sf_static_study sf_static_study;
typedef struct sf_csr sf_csr;
typedef struct sf_multigrid_setup sf_multigrid_setup;
typedef struct sf_multigrid sf_multigrid;
typedef struct sf_pcg sf_pcg;
typedef struct sf_pcg_batch sf_pcg_batch;
typedef struct sf_server sf_server;

/* ---- errors and device ---------------------------------------------------------------- */
const char *sf_last_error(void);
void sf_last_error_payload(int64_t *a, int64_t *b, double *value);
const char *sf_version(void);
sf_status sf_device_count(int32_t *count);
/* A device context: one GPU, one stream. No reference counterpart: the reference runs on the
 * host (crates/solidfem-solve/src/parallel.rs:82-215 is its worker pool). */
sf_status sf_context_create(int32_t device, sf_context **out);
void sf_context_destroy(sf_context *ctx);
sf_status sf_context_synchronize(sf_context *ctx);

/* ---- mesh: crates/solidfem-mesh ------------------------------------------------------- */
/* TetMesh::new, crates/solidfem-mesh/src/tet.rs:59-109 */
sf_status sf_tet_mesh_create(const double *positions, uint64_t node_count, const uint32_t *tets,
                             uint64_t tet_count, sf_tet_mesh **out);
/* A mesh from its text form, the on-wire form of a mesh request: `vertex-count tet-count`, then
 * `x y z` for every vertex, then `a b c d` (vertex indices from 0) for every tetrahedron, separated
 * by any white space -- the triangle-surface text of crates/solidfem-bench/src/bin/gate.rs:9-14
 * with tetrahedra in place of triangles (this path has no mesher to fill a surface). Validated as
 * TetMesh::new (tet.rs:59-109, SF_ERR_MESH_*); malformed text is refused with
 * SF_ERR_INVALID_ARGUMENT, a = the line. */
sf_status sf_tet_mesh_read_text(const char *text, uint64_t length, sf_tet_mesh **out);
/* box_tet_mesh, crates/solidfem-mesh/src/box_mesh.rs:22-80 */
sf_status sf_box_tet_mesh(const double origin[3], const double size[3], const int32_t divisions[3],
                          sf_tet_mesh **out);
void sf_tet_mesh_destroy(sf_tet_mesh *mesh);
uint64_t sf_tet_mesh_node_count(const sf_tet_mesh *mesh);
uint64_t sf_tet_mesh_tet_count(const sf_tet_mesh *mesh);
/* TetMesh::positions / tets, tet.rs:111-121: 3 per vertex, 4 per tetrahedron */
void sf_tet_mesh_positions(const sf_tet_mesh *mesh, double *out);
void sf_tet_mesh_tets(const sf_tet_mesh *mesh, uint32_t *out);
/* QuadraticTetMesh::from_linear, crates/solidfem-mesh/src/quadratic.rs:24-71 */
sf_status sf_quad_mesh_from_linear(const sf_tet_mesh *mesh, sf_quad_mesh **out);
void sf_quad_mesh_destroy(sf_quad_mesh *mesh);
uint64_t sf_quad_mesh_node_count(const sf_quad_mesh *mesh);
uint64_t sf_quad_mesh_tet_count(const sf_quad_mesh *mesh);
uint64_t sf_quad_mesh_vertex_count(const sf_quad_mesh *mesh);
void sf_quad_mesh_nodes(const sf_quad_mesh *mesh, double *out);   /* 3 per node */
void sf_quad_mesh_tets(const sf_quad_mesh *mesh, uint32_t *out);  /* 10 per element */
/* QuadraticTetMesh::boundary_faces, crates/solidfem-mesh/src/quadratic.rs:89-124:
 * (tet, face, 6 nodes) as 8 u32 per face. Returns the count; out may be NULL. */
uint64_t sf_quad_mesh_boundary_faces(const sf_quad_mesh *mesh, uint32_t *out);

/* ---- model and study: crates/solidfem-elastic ------------------------------------------ */
/* Model::new, crates/solidfem-elastic/src/study.rs:58-68 (uniform material) */
sf_status sf_model_create(const sf_quad_mesh *mesh, double youngs_modulus, double poisson_ratio,
                          sf_model **out);
/* A two-material field by depth below the long faces (the MaterialField of
 * crates/solidfem/tests/validation/shell_core.rs:19-21 over material.rs:35-50). */
sf_status sf_model_set_shell_core(sf_model *model, double shell_depth, double core_youngs,
                                  double core_poisson);
/* Model::set_quadrature, study.rs:75-78: 4 or 14 points */
sf_status sf_model_set_quadrature(sf_model *model, int32_t points);
/* Model::fix_node, study.rs:80-83 */
sf_status sf_model_fix_node(sf_model *model, uint32_t node);
/* Model::add_nodal_force, study.rs:85-88 */
sf_status sf_model_add_nodal_force(sf_model *model, uint32_t node, const double force[3]);
/* Model::prescribe_displacement, study.rs:90-99 */
sf_status sf_model_prescribe_displacement(sf_model *model, uint32_t node, uint32_t axis, double value);
/* Model::add_traction, study.rs:101-104; face = (tet, face, 6 nodes) as from boundary_faces */
sf_status sf_model_add_traction(sf_model *model, const uint32_t face[8], const double traction[3]);
/* A caller's MaterialField (material.rs:35-50): `at` writes the material (E, nu) at a physical
 * position. */
typedef void (*sf_material_fn)(const double position[3], double material[2], void *user);
/* Model::new with a caller's MaterialField: `at` is called once for every integration point of
 * the 4-point and then of the 14-point rule (element by element, points in rule order, at the
 * point's position as point_geometry forms it), and the model keeps the values (SF_MATERIAL_POINTS)
 * for its assembly, recovery and error estimate, whichever rule it solves with. Values are checked
 * where the reference checks them: an invalid material fails the assembly with
 * SF_ERR_MODEL_INVALID_MATERIAL. Replaces the model's material. */
sf_status sf_model_set_material_field(sf_model *model, sf_material_fn at, void *user);
/* A structured box model (mesh, clamp, tractions, material) from a workload description. */
sf_status sf_model_from_workload(const sf_workload *workload, sf_model **out);
void sf_model_destroy(sf_model *model);
const sf_quad_mesh *sf_model_mesh(const sf_model *model);
/* element_stiffness, crates/solidfem-elastic/src/element.rs:98-133, evaluated on the device for
 * elements [first, first+count): 900 values each, row-major. */
sf_status sf_model_element_stiffness(sf_context *ctx, const sf_model *model, uint64_t first,
                                     uint64_t count, double *ke);
/* assemble_stiffness, crates/solidfem-elastic/src/assembly.rs:17-27: the full stiffness with no
 * node eliminated, three unknowns per node in node order, assembled on the device (the study's
 * element kernels and scatter with every node free). It has the model's six rigid-body null
 * vectors. The model's constraints and loads are not read. The cancel check is read before each
 * chunk of 65536 elements (SF_ERR_MODEL_CANCELLED); an inverted element gives
 * SF_ERR_MODEL_INVERTED_JACOBIAN. The matrix stays on the device: sf_csr_read reads it back. */
sf_status sf_assemble_stiffness(sf_context *ctx, const sf_model *model, sf_cancel_fn cancel, void *user,
                                sf_csr **out);

/* ---- element functions (crates/solidfem-elastic/src/lib.rs:56-71), on the host, one point at a
 * time: the operations the device kernels apply to every element, for a caller's own element work.
 * Nodes and displacements are 10 x 3 in the Tet10 node order. */
/* shape_values, shape_derivatives (shape.rs:35-67) at natural (r, s, t): 10 values; 10 x 3
 * derivatives, out[3 a + axis] */
void sf_shape_values(const double natural[3], double out[10]);
void sf_shape_derivatives(const double natural[3], double out[30]);
/* Quadrature::points (quadrature.rs:44-77) of the 4- or 14-point rule: r, s, t and weight per
 * point into `points` (may be NULL); returns the count, 0 for another rule. */
uint64_t sf_quadrature_points(int32_t rule, double *points);
/* PointGeometry (element.rs:14-23): physical shape gradients (10 x 3), det J, position */
typedef struct sf_geometry {
    double gradients[30];
    double det_j;
    double position[3];
} sf_geometry;
/* point_geometry (element.rs:25-80): SF_ERR_MODEL_INVERTED_JACOBIAN (det_j and position still
 * written) when det J is not positive or NaN */
sf_status sf_point_geometry(const double nodes[30], const double natural[3], sf_geometry *out);
/* strain_at (element.rs:160-171): [xx, yy, zz, 2xy, 2yz, 2zx] */
void sf_strain_at(const sf_geometry *geometry, const double displacements[30], double strain[6]);
/* stress_at (element.rs:174-186): [xx, yy, zz, xy, yz, zx] of an isotropic material */
void sf_stress_at(double youngs_modulus, double poisson_ratio, const double strain[6], double stress[6]);
/* von_mises (element.rs:189-192) */
double sf_von_mises(const double stress[6]);

/* The numbered problem of StudyBuilder::new (study.rs:436-449, assembly.rs:55-81) on the host:
 * free map (node -> free index or UINT32_MAX), restricted right-hand side, the free stiffness's
 * block pattern (block row starts, block columns) and rigid-body modes (rigid_modes.rs:36-83).
 * Any output may be NULL; the counts are returned. */
sf_status sf_model_free_problem(const sf_model *model, uint64_t *free_nodes, uint64_t *blocks, uint32_t *free_map,
                                double *rhs, int64_t *block_starts, uint32_t *block_columns, double *rigid_modes);

/* StudyBuilder::new, crates/solidfem-elastic/src/study.rs:419-474. use_multigrid=0 chooses
 * Jacobi (PreconditionerChoice::Jacobi, study.rs:108-117); mg may be NULL then. The model is
 * copied. */
sf_status sf_study_builder_create(sf_context *ctx, const sf_model *model, const sf_pcg_options *pcg,
                                  int32_t use_multigrid, const sf_multigrid_options *mg,
                                  sf_study_builder **out);
/* StudyBuilder::step, study.rs:478-512 */
sf_status sf_study_builder_step(sf_study_builder *builder, uint64_t max_elements, sf_cancel_fn cancel,
                                void *user, sf_build_progress *progress);
/* StudyBuilder::finish, study.rs:601-656; consumes the builder on success */
sf_status sf_study_builder_finish(sf_study_builder *builder, sf_static_study **out);
void sf_study_builder_destroy(sf_study_builder *builder);
/* StaticStudy::new, study.rs:191-199 (builder step with an unbounded budget, then finish) */
sf_status sf_static_study_create(sf_context *ctx, const sf_model *model, const sf_pcg_options *pcg,
                                 int32_t use_multigrid, const sf_multigrid_options *mg,
                                 sf_cancel_fn cancel, void *user, sf_static_study **out);
/* StaticStudy::step, study.rs:202-215 */
sf_status sf_static_study_step(sf_static_study *study, uint64_t max_iterations, sf_cancel_fn cancel,
                               void *user, sf_progress *progress);
/* StaticStudy::displacements, study.rs:218-237: 3 per mesh node, fixed nodes zero */
sf_status sf_static_study_displacements(const sf_static_study *study, double *out);
/* StaticStudy::report, study.rs:240-243 */
sf_status sf_static_study_report(const sf_static_study *study, sf_report *out);
/* StaticStudy::multigrid, study.rs:362-368 (NULL for Jacobi) */
const sf_multigrid *sf_static_study_multigrid(const sf_static_study *study);
/* The conjugate gradient solve behind the study. */
sf_pcg *sf_static_study_pcg(sf_static_study *study);
/* The assembled free stiffness (scalar CSR, the reference's CsrMatrix, csr.rs:6-13), read back
 * from the device: row_starts (n+1), columns and values (nnz); any pointer may be NULL. */
sf_status sf_static_study_matrix(const sf_static_study *study, uint64_t *size, uint64_t *nonzeros,
                                 int64_t *row_starts, uint32_t *columns, double *values);
sf_status sf_static_study_rhs(const sf_static_study *study, double *out);
void sf_static_study_destroy(sf_static_study *study);

/* ---- linear algebra: crates/solidfem-solve --------------------------------------------- */
/* CsrMatrix from plain arrays (the pattern of CsrMatrix::from_block_connectivity plus values,
 * csr.rs:19-71 and :89-96), uploaded to the device; with ctx NULL it stays on the host (enough
 * for a multigrid setup, which runs on the host). */
sf_status sf_csr_create(sf_context *ctx, uint64_t size, const int64_t *row_starts,
                        const uint32_t *columns, const double *values, sf_csr **out);
void sf_csr_destroy(sf_csr *csr);
/* The matrix as scalar CSR (CsrMatrix's row_starts, columns and values, csr.rs:6-13): size,
 * nonzeros, row_starts (size+1), columns and values (nonzeros); any pointer may be NULL. A matrix
 * on the device is read from there. */
sf_status sf_csr_read(const sf_csr *csr, uint64_t *size, uint64_t *nonzeros, int64_t *row_starts,
                      uint32_t *columns, double *values);
/* CsrMatrix::multiply, csr.rs:112-132, on the device: host x and y of `size` */
sf_status sf_csr_multiply(sf_csr *csr, const double *x, double *y);
/* CsrMatrix::multiply_rounding_bound, csr.rs:142-154, on the device */
sf_status sf_csr_rounding_bound(sf_csr *csr, const double *x, int32_t dot_order, double *bound);

/* The rounding bound of sf_csr_rounding_bound, allocating nothing. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). x (size values) and bound (one value) are
 * device pointers; the call enqueues its kernels on `stream` (a hipStream_t; NULL = the matrix's
 * context stream) and returns without waiting; *bound then holds the same bits as
 * sf_csr_rounding_bound gives for the same x and dot order. Its scratch is one 16 MiB device array
 * per GPU, set up when a context is created on that GPU and kept until the process ends, so a
 * matrix of more than sf_noalloc_max_rows() scalar rows is refused (SF_ERR_INVALID_ARGUMENT,
 * a = rows, b = the limit). Every call on a GPU shares that array: calls that overlap on one GPU
 * (on different streams, or from different threads) race and must be serialised by the caller. */
sf_status sf_csr_rounding_bound_noalloc(const sf_csr *csr, const double *x, int32_t dot_order, double *bound, void *stream);
uint64_t sf_noalloc_max_rows(void);

/* MultigridSetup::new, crates/solidfem-solve/src/multigrid/setup.rs:129-189: near_null holds
 * near_null_len values (6 per scalar row), positions 3 per block row; wrong lengths are refused. */
sf_status sf_multigrid_setup_create(sf_context *ctx, const sf_csr *matrix, const double *near_null,
                                    uint64_t near_null_len, const double *positions, uint64_t position_count,
                                    const sf_multigrid_options *options, sf_multigrid_setup **out);
/* MultigridSetup::step, setup.rs:193-222: finished = 1 once done */
sf_status sf_multigrid_setup_step(sf_multigrid_setup *setup, uint64_t budget, sf_cancel_fn cancel,
                                  void *user, int32_t *finished);
/* MultigridSetup::finish, setup.rs:230-240: builds the device hierarchy */
sf_status sf_multigrid_setup_finish(sf_multigrid_setup *setup, sf_multigrid **out);
void sf_multigrid_setup_destroy(sf_multigrid_setup *setup);
void sf_multigrid_destroy(sf_multigrid *mg);
/* Multigrid::stats / precision, cycle.rs:319-330 */
sf_status sf_multigrid_stats_get(const sf_multigrid *mg, sf_multigrid_stats *out);
int32_t sf_multigrid_precision(const sf_multigrid *mg);
/* Multigrid::threads (cycle.rs:310-316): the host threads the cycle's products run on -- one, the
 * calling thread, which enqueues them on the device (the bits are the same as with any count). */
uint64_t sf_multigrid_threads(const sf_multigrid *mg);
/* Preconditioner::apply for Multigrid, cycle.rs:430-446: one V-cycle on the device, host arrays */
sf_status sf_multigrid_apply(sf_multigrid *mg, const double *residual, double *output);
/* One Chebyshev smoother application on `level` from a zero guess, cycle.rs:570-651 */
sf_status sf_multigrid_smooth(sf_multigrid *mg, uint64_t level, const double *b, double *x);
/* Multigrid::level_sizes, cycle.rs:332-339: (rows, row_block, stored values) per level */
sf_status sf_multigrid_level_sizes(const sf_multigrid *mg, uint64_t *out, uint64_t capacity, uint64_t *levels);
/* Multigrid::level_matrix / level_prolongator / level_inverse_diagonal, cycle.rs:363-395, values
 * widened to f64 as stored on the device. Any output pointer may be NULL; counts are returned. */
sf_status sf_multigrid_level_matrix(const sf_multigrid *mg, uint64_t level, int32_t prolongator,
                                    uint64_t *rows, uint64_t *block_columns, uint64_t *row_block,
                                    uint64_t *col_block, uint64_t *blocks, int64_t *starts,
                                    uint32_t *columns, double *values);
sf_status sf_multigrid_level_inverse_diagonal(const sf_multigrid *mg, uint64_t level, double *out);
/* Multigrid::smoother_spectrum / aggregates, cycle.rs:341-361 */
sf_status sf_multigrid_smoother_spectrum(const sf_multigrid *mg, uint64_t level, double *lambda_min,
                                         double *lambda_max);
sf_status sf_multigrid_aggregates(const sf_multigrid *mg, uint64_t level, uint32_t *out, uint64_t *count);
/* Multigrid::solve_coarsest, cycle.rs:402-408 */
sf_status sf_multigrid_solve_coarsest(sf_multigrid *mg, double *values, uint64_t count);

/* The levels of a finished setup, before MultigridSetup::finish moves them to the device: the
 * same views as the sf_multigrid_level_* calls, in f64 as set up (host only, no device needed). */
sf_status sf_multigrid_setup_level_count(const sf_multigrid_setup *setup, uint64_t *levels);
sf_status sf_multigrid_setup_level_matrix(const sf_multigrid_setup *setup, uint64_t level, int32_t prolongator,
                                          uint64_t *rows, uint64_t *block_columns, uint64_t *row_block,
                                          uint64_t *col_block, uint64_t *blocks, int64_t *starts,
                                          uint32_t *columns, double *values);
sf_status sf_multigrid_setup_level_inverse_diagonal(const sf_multigrid_setup *setup, uint64_t level, double *out);
sf_status sf_multigrid_setup_smoother_spectrum(const sf_multigrid_setup *setup, uint64_t level, double *lambda_min,
                                               double *lambda_max);
sf_status sf_multigrid_setup_aggregates(const sf_multigrid_setup *setup, uint64_t level, uint32_t *out, uint64_t *count);
/* The coarsest level's Cholesky factor (n x n row-major, lower triangle); *n = 0 when unfactored. */
sf_status sf_multigrid_setup_coarse_factor(const sf_multigrid_setup *setup, uint64_t *n, double *out);

/* Pcg::new, crates/solidfem-solve/src/pcg.rs:135-164, preconditioned by a multigrid (mg) or,
 * with mg NULL, by Jacobi (preconditioner.rs:53-66). The solve owns device copies; the matrix
 * and the multigrid must outlive it. */
sf_status sf_pcg_create(sf_context *ctx, const sf_csr *matrix, sf_multigrid *mg, const double *rhs,
                        const sf_pcg_options *options, sf_pcg **out);
/* The Preconditioner seam (preconditioner.rs:9-20): a caller's M^-1, applied by a callback that
 * writes z = M^-1 r for the n unknowns and returns 0 (non-zero fails the step with
 * SF_ERR_PRECONDITIONER, a = the value). It must be linear, symmetric and positive definite.
 *   SF_APPLY_HOST: r and z are host arrays (Preconditioner::apply's slices); the library copies
 *     r out of HBM before the call and z back after it, 2 n doubles over PCIe per iteration.
 *   SF_APPLY_DEVICE: r and z are device pointers; the callback enqueues its work on `stream`
 *     (the solve's hipStream_t) and returns; the solve's next launch follows it on that stream.
 * The callback runs on the thread that steps the solve. */
typedef int32_t (*sf_apply_fn)(const double *r, double *z, uint64_t n, void *stream, void *user);
enum { SF_APPLY_HOST = 0, SF_APPLY_DEVICE = 1 };
typedef struct sf_preconditioner {
    sf_apply_fn apply;
    void *user;
    int32_t where; /* SF_APPLY_HOST or SF_APPLY_DEVICE */
    int32_t reserved;
} sf_preconditioner;
/* Pcg::new with a caller's preconditioner; the callback and its user data must outlive the solve. */
sf_status sf_pcg_create_with(sf_context *ctx, const sf_csr *matrix, const sf_preconditioner *preconditioner,
                             const double *rhs, const sf_pcg_options *options, sf_pcg **out);
/* Pcg::step, pcg.rs:168-189 */
sf_status sf_pcg_step(sf_pcg *pcg, uint64_t max_iterations, sf_cancel_fn cancel, void *user,
                      sf_progress *progress);
/* Pcg::estimate, pcg.rs:192-194 (copied to the host) */
sf_status sf_pcg_estimate(const sf_pcg *pcg, double *out);
/* Pcg::progress / is_converged, pcg.rs:202-213 */
sf_status sf_pcg_progress(const sf_pcg *pcg, sf_progress *out);
int32_t sf_pcg_is_converged(const sf_pcg *pcg);
uint64_t sf_pcg_size(const sf_pcg *pcg);
void sf_pcg_destroy(sf_pcg *pcg);

/* ---- resident re-use and measurement (no reference counterpart) ------------------------ */
/* The study's system matrix as a CsrMatrix handle (owned by the study). */
const sf_csr *sf_static_study_system(const sf_static_study *study);
/* The same hierarchy stored and applied in another precision (Multigrid::from_levels,
 * crates/solidfem-solve/src/multigrid/cycle.rs:265-308, over the same f64 levels), with the
 * given coarse solve (SF_COARSE_*, or -1 to keep this one's). */
sf_status sf_multigrid_with_precision(const sf_multigrid *mg, int32_t precision, int32_t coarse_solve,
                                      sf_multigrid **out);
/* Launch option, no reference counterpart: with enable set, sf_multigrid_apply and the solves
 * using this hierarchy replay a HIP graph of one V-cycle's launches, captured on first use for
 * each pair of input and output buffers, instead of launching the kernels one by one. Results are
 * the same bits. Off by default; never used while kernel timing (sf_profile_enable) is on.
 * Enabling it is refused with SF_ERR_GRAPH_UNDER_TRACER while rocprofv3 traces the process. */
sf_status sf_multigrid_set_graph(sf_multigrid *mg, int32_t enable);
/* Algorithmic HBM bytes of one V-cycle of this hierarchy (every array a kernel touches, once). */
double sf_multigrid_cycle_bytes(const sf_multigrid *mg);
/* Restarts a solve from a zero estimate (Pcg::new semantics, pcg.rs:135-164) on the buffers it
 * already holds; new_rhs NULL keeps the right-hand side. Allocates nothing. */
sf_status sf_pcg_restart(sf_pcg *pcg, const double *new_rhs);
/* ---- stress recovery -------------------------------------------------------------------------
 * Recovery (crates/solidfem-elastic/src/results.rs:307-576 with element.rs:135-192 and patch.rs),
 * on the device, for nodal displacements (3 per mesh node, fixed components included):
 *   point_values: per element and integration point, in (element, point) order, 10 values --
 *     position (3), stress [xx, yy, zz, xy, yz, zx] (6), von Mises (1) -- PointStress;
 *   nodal_values: per mesh node 7 values -- the recovered nodal stress (6), from the vertex patch
 *     fits, midside means and the worst incident value where a patch is not fitted, and its von
 *     Mises (1) -- Solution::nodal_stress and nodal_von_mises.
 * Either output may be NULL. The first element that fails, in element then point order, fails the
 * call: SF_ERR_MODEL_INVERTED_JACOBIAN (a = the element, b = the integration point, or the node
 * for the node values) or SF_ERR_MODEL_INVALID_MATERIAL (a = the element, b = the point). */
sf_status sf_recover_stress(sf_context *ctx, const sf_model *model, const double *displacements, double *point_values,
                            double *nodal_values);
/* The same for a study's current displacements (Recovery as StaticStudy drives it, study.rs). */
sf_status sf_static_study_recover(sf_static_study *study, double *point_values, double *nodal_values);
/* The resumable form (Recovery, results.rs:103-124 and :440-536): a snapshot of displacements
 * recovered in slices -- every element in order, then one patch fit per mesh vertex in order --
 * whose finished field has the bits of sf_recover_stress whatever the slicing. The recovery
 * holds its own device copies of what it reads and outlives the model or study it came from. */
typedef struct sf_recovery sf_recovery;
typedef struct sf_recovery_state {
    uint64_t elements, elements_done; /* Recovery::elements_done */
    uint64_t vertices, fits_done;     /* patch fits, one per mesh vertex */
    uint64_t nodes, points_per_element;
    int32_t converged;                /* the displacements are a converged solve's */
    int32_t finished;
} sf_recovery_state;
/* A recovery of `displacements` (3 per mesh node) over `model`, none of it done. */
sf_status sf_recovery_create(sf_context *ctx, const sf_model *model, const double *displacements, int32_t converged,
                             sf_recovery **out);
/* StaticStudy::begin_recovery (study.rs:263-273): the study's current displacements, the converged
 * answer or a preview of a solve in progress. */
sf_status sf_static_study_begin_recovery(sf_static_study *study, sf_recovery **out);
/* Recovery::advance (results.rs:477-536): up to max_work more units, an element, then a patch fit;
 * fits start only once every element is done. `cancel` (may be NULL) is read before each unit:
 * SF_ERR_MODEL_CANCELLED, with the units before it done. A faulting element fails as
 * sf_recover_stress says (a = the element), and stays the next to do. `finished`
 * (may be NULL) is set to whether every unit is done. */
sf_status sf_recovery_advance(sf_recovery *recovery, uint64_t max_work, sf_cancel_fn cancel, void *user, int32_t *finished);
/* StaticStudy::continue_recovery (study.rs:275-292): sf_recovery_advance, its time added to the
 * study's recovery_seconds. */
sf_status sf_static_study_continue_recovery(sf_static_study *study, sf_recovery *recovery, uint64_t max_work,
                                            sf_cancel_fn cancel, void *user, int32_t *finished);
sf_status sf_recovery_state_of(const sf_recovery *recovery, sf_recovery_state *out);
/* The displacements being recovered, 3 per mesh node (Recovery::displacements). */
sf_status sf_recovery_displacements(const sf_recovery *recovery, double *out);
/* The finished field in sf_recover_stress's layout (either output may be NULL);
 * SF_ERR_MODEL_NOT_FINISHED (a = elements done, b = fits done) while units are left
 * (StaticStudy::finish_recovery, study.rs:294-300). */
sf_status sf_recovery_field(sf_recovery *recovery, double *point_values, double *nodal_values);
void sf_recovery_destroy(sf_recovery *recovery);

/* ---- error estimate (crates/solidfem-elastic/src/error_estimate.rs) --------------------------
 * ErrorEstimation: how converged a finished recovery's solution is -- per element the energy norm
 * e_K of the recovered minus the element stress, integrated with the 14-point rule whatever rule
 * the solve used, and over the model ||e||^2 = sum e_K^2, ||u_h||^2 (twice the strain energy) and
 * eta = ||e|| / sqrt(||u_h||^2 + ||e||^2). The element integrals run on the device; the sums are
 * kept in element order, so any slicing gives one run's bits. The estimation borrows the recovery
 * (ErrorEstimation holds it, error_estimate.rs:80-89): destroy the estimation first. */
typedef struct sf_error_estimation sf_error_estimation;
typedef struct sf_error_estimate {
    double error_norm_squared;    /* ||e||^2 */
    double solution_norm_squared; /* ||u_h||^2 */
    double relative_error;        /* eta, in [0, 1) */
    uint64_t elements;            /* element_errors' length */
} sf_error_estimate;
/* StaticStudy::begin_error_estimate (study.rs:329-333): an estimate of `recovery` over the study's
 * mesh and material, none of it done. */
sf_status sf_static_study_begin_error_estimate(const sf_static_study *study, sf_recovery *recovery,
                                               sf_error_estimation **out);
/* The same over a model's mesh and material. */
sf_status sf_error_estimation_create(const sf_model *model, sf_recovery *recovery, sf_error_estimation **out);
/* ErrorEstimation::advance (error_estimate.rs:121-151): up to max_elements more elements in order,
 * `cancel` (may be NULL) read before each (SF_ERR_MODEL_CANCELLED, the elements before it done).
 * Refuses a recovery of another mesh (SF_ERR_MODEL_RECOVERY_MESH_MISMATCH) and an unfinished or
 * preview recovery (SF_ERR_MODEL_NOT_FINISHED). An element whose corner has no patch fit gives
 * SF_ERR_MODEL_MIXED_MATERIAL_PATCH or SF_ERR_MODEL_COPLANAR_PATCH (a = vertex); an inverted point
 * SF_ERR_MODEL_INVERTED_JACOBIAN, an invalid material SF_ERR_MODEL_INVALID_MATERIAL (a = element,
 * b = point of the 14-point rule); the elements before it are done. `finished` may be NULL. */
sf_status sf_error_estimation_advance(sf_error_estimation *estimation, uint64_t max_elements, sf_cancel_fn cancel,
                                      void *user, int32_t *finished);
/* StaticStudy::continue_error_estimate (study.rs:335-352): sf_error_estimation_advance, its time
 * added to the study's recovery_seconds. */
sf_status sf_static_study_continue_error_estimate(sf_static_study *study, sf_error_estimation *estimation,
                                                  uint64_t max_elements, sf_cancel_fn cancel, void *user,
                                                  int32_t *finished);
/* ErrorEstimation::elements_done */
uint64_t sf_error_estimation_elements_done(const sf_error_estimation *estimation);
/* ErrorEstimation::estimate (error_estimate.rs:153-171), StaticStudy::finish_error_estimate:
 * element_errors (one per element, may be NULL) and the sums. SF_ERR_MODEL_NOT_FINISHED with
 * elements left, SF_ERR_MODEL_NON_FINITE_ENERGY, SF_ERR_MODEL_ZERO_STRAIN_ENERGY. */
sf_status sf_error_estimation_finish(const sf_error_estimation *estimation, sf_error_estimate *out,
                                     double *element_errors);
void sf_error_estimation_destroy(sf_error_estimation *estimation);

/* ---- peak stress (results.rs:149-226) ---------------------------------------------------------
 * The peak von Mises over the positions `inside` accepts (NULL: everywhere): the mesh nodes in
 * order, then the integration points in order, each classified by its own physical position. A
 * later candidate replaces the peak only when strictly larger, so the first of equal values holds
 * it; the first NaN met becomes the peak and stays it. node_positions: 3 per node; the nodal von
 * Mises of node n at nodal_von_mises[n * nodal_stride] (7 with sf_recover_stress's nodal values,
 * reading from their von Mises column); point_values: 10 per point as sf_recover_stress writes
 * them. peak_von_mises is this peak's value. */
typedef int32_t (*sf_region_fn)(const double *position, void *user);
typedef struct sf_peak {
    int32_t found;     /* 0: no node and no point is inside */
    int32_t at_point;  /* held by an integration point (else a mesh node): PeakHolder */
    uint64_t index;    /* of that node or point */
    double von_mises;
    double position[3];
} sf_peak;
sf_status sf_peak_stress(const double *node_positions, uint64_t nodes, const double *nodal_von_mises, uint64_t nodal_stride,
                         const double *point_values, uint64_t points, sf_region_fn inside, void *user, sf_peak *out);

/* ---- mesh comparison (crates/solidfem-elastic/src/comparison.rs) ------------------------------
 * How much the peak von Mises moved between a coarser and a finer solve of one part and load, read
 * over a region (the region callback, NULL: everywhere). A solve is given as its mesh, its nodal
 * von Mises (`nodes` values, `nodal_stride` apart: 7 with sf_recover_stress's nodal values, from
 * its von Mises column) and its point values (10 per point as sf_recover_stress writes them, the
 * same number per element in element order). Host work over finished solutions. */
typedef struct sf_compared_solve {
    const sf_quad_mesh *mesh;
    const double *nodal_von_mises;
    uint64_t nodes, nodal_stride;
    const double *point_values;
    uint64_t points;
} sf_compared_solve;
/* ComparedPeak (comparison.rs:55-66) */
typedef struct sf_compared_peak {
    sf_peak peak;
    uint64_t elements;          /* elements with an integration point inside the region */
    double peak_element_volume; /* the peak's element, or the mean of the elements holding its node */
} sf_compared_peak;
/* MeshComparison (comparison.rs:83-92) */
typedef struct sf_mesh_comparison {
    sf_compared_peak coarse, fine;
    double peak_change;   /* (fine - coarse) / fine of the peak */
    double peak_distance; /* between the two peaks' positions */
} sf_mesh_comparison;
/* Verdict (comparison.rs:68-80) */
enum { SF_VERDICT_SETTLED = 0, SF_VERDICT_STILL_MOVING = 1, SF_VERDICT_MESHES_TOO_ALIKE = 2 };
/* MIN_ELEMENT_RATIO (comparison.rs:53) */
#define SF_MIN_ELEMENT_RATIO 4.0
/* MeshComparison::between (comparison.rs:101-141): SF_ERR_MODEL_SOLUTION_MESH_MISMATCH,
 * SF_ERR_MODEL_NOTHING_INSIDE_REGION, SF_ERR_MODEL_NON_FINITE_PEAK, SF_ERR_MODEL_ZERO_PEAK (a = the
 * solve), SF_ERR_MODEL_FINE_MESH_NOT_FINER (a, b = the element counts). The region
 * callback returns nonzero for a position inside. */
sf_status sf_mesh_comparison_between(const sf_compared_solve *coarse, const sf_compared_solve *fine, sf_region_fn inside,
                                     void *user, sf_mesh_comparison *out);
/* element_ratio, peak_element_ratio (comparison.rs:154-169) */
double sf_mesh_comparison_element_ratio(const sf_mesh_comparison *comparison);
double sf_mesh_comparison_peak_element_ratio(const sf_mesh_comparison *comparison);
/* safety_factors and safety_factor_change (comparison.rs:186-205): SF_ERR_MODEL_INVALID_YIELD_STRENGTH
 * for a yield strength that is not finite and positive */
sf_status sf_mesh_comparison_safety_factors(const sf_mesh_comparison *comparison, double yield_strength, double *coarse,
                                            double *fine);
sf_status sf_mesh_comparison_safety_factor_change(const sf_mesh_comparison *comparison, double yield_strength,
                                                  double *change);
/* verdict (comparison.rs:207-222): SF_VERDICT_* */
int32_t sf_mesh_comparison_verdict(const sf_mesh_comparison *comparison, double tolerance);
/* nodal_rotations (results.rs:262-305): the local rotation (half the curl of the displacement), in
 * radians, at every mesh node -- each incident element's value at the node weighted by its
 * Jacobian determinant there, summed in element order. rotations: 3 per node. */
sf_status sf_nodal_rotations(sf_context *ctx, const sf_model *model, const double *displacements, double *rotations);
/* A study's mesh nodes, elements and integration points per element. */
sf_status sf_static_study_counts(const sf_static_study *study, uint64_t *nodes, uint64_t *elements, uint64_t *points);
/* Integration points per element of a model's rule (4 or 14). */
uint64_t sf_model_points_per_element(const sf_model *model);

/* ---- several load cases over one matrix and one hierarchy ------------------------------------
 * No reference counterpart: the reference solves one right-hand side per Pcg
 * (crates/solidfem-solve/src/pcg.rs:113-283). A batch runs that recurrence for up to 64
 * right-hand sides in lockstep, each column on its own scalars, so that every pass over the
 * matrix and the hierarchy serves all of them. Column m's iterates, iteration count and outcome
 * are bit for bit those of sf_pcg_create with column m alone (same options), whatever the other
 * columns hold. Right-hand sides and estimates are column-major: column m at [m * n, (m + 1) * n). */
#define SF_MAX_COLUMNS 64
/* Z = M^-1 R for `columns` columns (column-major in and out); column m equals
 * sf_multigrid_apply of column m bit for bit. */
sf_status sf_multigrid_apply_batch(sf_multigrid *mg, const double *residuals, uint64_t columns, double *outputs);
sf_status sf_pcg_batch_create(sf_context *ctx, const sf_csr *matrix, sf_multigrid *mg, const double *rhs,
                              uint64_t columns, const sf_pcg_options *options, sf_pcg_batch **out);
/* Up to max_iterations lockstep iterations, the cancel callback polled before each (it cancels
 * every running column). `progress` receives one entry per column; `running` the columns still
 * running. Column failures are reported per column, not as the call's status. */
sf_status sf_pcg_batch_step(sf_pcg_batch *batch, uint64_t max_iterations, sf_cancel_fn cancel, void *user,
                            sf_column_progress *progress, uint64_t *running);
sf_status sf_pcg_batch_progress(const sf_pcg_batch *batch, sf_column_progress *progress);
sf_status sf_pcg_batch_estimate(const sf_pcg_batch *batch, double *out);
/* Restarts every column from a zero estimate; rhs NULL keeps the right-hand sides. */
sf_status sf_pcg_batch_restart(sf_pcg_batch *batch, const double *rhs);
uint64_t sf_pcg_batch_columns(const sf_pcg_batch *batch);
void sf_pcg_batch_destroy(sf_pcg_batch *batch);

/* ---- serving ------------------------------------------------------------------------------
 * No reference counterpart: the reference solves one study at a time on the caller's thread
 * (crates/solidfem-elastic/src/study.rs:191-372). A server takes structured box problems
 * (sf_workload) from any thread and answers each with its displacements. Requests with the same
 * mesh, constraints and material share one resident system -- the assembled matrix and its
 * multigrid hierarchy -- kept in HBM under a byte budget (least recently used idle systems are
 * evicted to admit new ones; a system that cannot fit even alone is refused). A system's requests
 * are solved as the columns of one lockstep batch (sf_pcg_batch) and new requests take free
 * columns between slices of iterations, so every answer is bit for bit the answer the request
 * gets alone. Latency is measured from submission to the answer. */
enum { SF_SERVE_QUEUED = 0, SF_SERVE_BUILDING = 1, SF_SERVE_SOLVING = 2, SF_SERVE_DONE = 3, SF_SERVE_FAILED = 4 };
typedef struct sf_server_options {
    uint64_t hbm_budget;       /* bytes the server may reserve; 0 = 200 GB */
    int32_t precision;         /* SF_PRECISION_* of the V-cycle */
    int32_t coarse_solve;      /* SF_COARSE_* */
    double tolerance;          /* PcgOptions::tolerance, pcg.rs:14-33 */
    uint32_t max_columns;      /* requests per batch, 1 to 64 */
    uint32_t slice_iterations; /* lockstep iterations between admissions of new requests */
    int32_t builders;          /* threads assembling systems and setting up hierarchies */
    int32_t solvers;           /* threads running batches, one stream each */
    int32_t dot_order;         /* SF_DOT_* */
    int32_t reserved;
} sf_server_options;
typedef struct sf_serve_status {
    int32_t state;             /* SF_SERVE_* */
    sf_status error;           /* for SF_SERVE_FAILED */
    int32_t cache_hit;         /* the request found its system resident or being built */
    int32_t at_rounding_floor;
    uint64_t iterations;
    double relative_residual;
    double queued_seconds;     /* waiting for its system or for a column */
    double build_seconds;      /* building the system this request brought */
    double solve_seconds;
    double latency_seconds;    /* submission to answer */
    uint64_t nodes;
    uint64_t free_dofs;
} sf_serve_status;
typedef struct sf_server_stats {
    uint64_t submitted, completed, failed, builds, cache_hits, evictions, resident_systems;
    uint64_t hbm_reserved, hbm_peak;
    double p50_seconds, p99_seconds, mean_seconds, max_seconds; /* latency of completed requests */
} sf_server_stats;
sf_status sf_server_create(int32_t device, const sf_server_options *options, sf_server **out);
sf_status sf_server_submit(sf_server *server, const sf_workload *workload, uint64_t *id);
/* A model of any mesh (copied): requests whose models have the same mesh, fixed nodes, prescribed
 * components (with their values), material and rule share one system; each brings its own nodal
 * forces and tractions, and its answer is that model's StaticStudy answer with the server's
 * options, bit for bit. */
sf_status sf_server_submit_model(sf_server *server, const sf_model *model, uint64_t *id);
sf_status sf_server_status(sf_server *server, uint64_t id, sf_serve_status *out);
/* Waits until the request is done or failed, or timeout_seconds pass (negative: no limit). */
sf_status sf_server_wait(sf_server *server, uint64_t id, double timeout_seconds, sf_serve_status *out);
/* The displacements of a finished request: 3 per mesh node (len), fixed components zero. */
sf_status sf_server_result(sf_server *server, uint64_t id, double *displacements, uint64_t len);
/* Forgets a finished request. */
sf_status sf_server_release(sf_server *server, uint64_t id);
sf_status sf_server_stats_get(sf_server *server, sf_server_stats *out);
/* Stops the threads; requests still open fail. */
void sf_server_destroy(sf_server *server);

/* The batch depth (blocks in flight per lane, 1 to 8) the tuning chose for a level's fused
 * Chebyshev step (kind 0) or residual product (kind 1), and for a matrix's conjugate-gradient
 * product A p. No reference counterpart: measurement reports which variant it timed. */
sf_status sf_multigrid_tuned_depth(const sf_multigrid *mg, uint64_t level, int32_t kind, int32_t *depth);
sf_status sf_csr_tuned_depth(const sf_csr *csr, int32_t *depth);

/* Tuning and fallback hooks, no reference counterpart (for tests and measurement). Both replace
 * environment variables, which a caller could only change with setenv while library threads may
 * read the environment (undefined behaviour in glibc); the environment is read once, as the
 * starting value.
 * sf_force_variant: every later tuning of a batch depth or batched form that has `code` among its
 *   candidates takes it (0: time the candidates, the default; SOLIDFEM_BATCH). Bits never depend
 *   on the variant.
 * sf_setup_slot_limit: device setup rows of more than `blocks` output blocks take the host path
 *   (0: the kernels' own limit of 1024; SOLIDFEM_SETUP_SLOT_LIMIT). Read when a setup begins. */
sf_status sf_force_variant(int32_t code);
sf_status sf_setup_slot_limit(uint32_t blocks);

/* Kernel launch timing with HIP events on the launching stream. Kernels: 0 = the fused
 * Chebyshev step on the finest level, 1 = the conjugate gradient's system product A p, 2 = the
 * residual on the finest level; for several columns, 3 = the finest level's batched Chebyshev
 * steps and 4 = the batched system product. read returns launches, summed seconds and summed
 * algorithmic bytes since the last reset. */
enum { SF_KERNEL_FINE_CHEBYSHEV = 0, SF_KERNEL_SYSTEM_PRODUCT = 1, SF_KERNEL_FINE_RESIDUAL = 2,
       SF_KERNEL_BATCH_FINE_CHEBYSHEV = 3, SF_KERNEL_BATCH_SYSTEM_PRODUCT = 4 };
sf_status sf_profile_enable(int32_t kernel, int32_t enable);
sf_status sf_profile_reset(void);
sf_status sf_profile_read(int32_t kernel, uint64_t *launches, double *seconds, double *bytes);
/* A plain device copy of `bytes` bytes (16-byte vector loads and stores), run `repeats` times on
 * the context's stream: the achievable HBM bandwidth and the calibration of byte counters.
 * Returns the average seconds per copy, timed with HIP events. */
sf_status sf_stream_copy(sf_context *ctx, uint64_t bytes, uint32_t repeats, double *seconds);
/* The same for a plain read of `bytes` bytes: the achievable HBM read bandwidth. */
sf_status sf_stream_read(sf_context *ctx, uint64_t bytes, uint32_t repeats, double *seconds);

#ifdef __cplusplus
}
#endif
#endif
