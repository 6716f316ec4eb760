/*
 * refcpu: the CPU restatement of solidfem's linear-static solve path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the oracle the HIP product is checked
 * against. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker or the timed CPU baseline -- never as part of the product.
 *
 * Every function restates the reference algorithm in the same floating-point order
 * (IEEE f64/f32, no fused multiply-add: built with -ffp-contract=off). It is pinned by
 * the reference's own golden fingerprints
 * (crates/solidfem/tests/validation/multigrid_precision.rs:134-250).
 *
 * A "workload" is a structured box problem: the box mesh of
 * crates/solidfem-mesh/src/box_mesh.rs:22-80 upgraded to Tet10
 * (crates/solidfem-mesh/src/quadratic.rs:24-71), clamped planes, face tractions and
 * an isotropic material field.
 */
#ifndef SOLIDFEM_REFCPU_H
#define SOLIDFEM_REFCPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Traction kinds applied to the boundary faces lying on a plane. */
enum ref_traction_kind {
    REF_TRACTION_UNIFORM = 0, /* the same traction vector on every selected face */
    REF_TRACTION_RANDOM = 1,  /* per face, each component N(0, scale) from (seed, case, face) */
    REF_TRACTION_TORSION = 2  /* scale * e_theta about the axis through (axis_u, axis_v) */
};

/* Material field kinds. */
enum ref_material_kind {
    REF_MATERIAL_UNIFORM = 0,
    REF_MATERIAL_SHELL_CORE = 1 /* depth to the nearest long face below shell_depth: shell */
};

typedef struct ref_workload {
    double origin[3];
    double size[3];
    int32_t divisions[3];
    /* interior-vertex jitter: fraction of the cell size, 0 for none */
    double jitter_fraction;
    uint64_t jitter_seed;
    /* clamp: nodes with |x[clamp_axis] - clamp_value| < 1e-9 are fixed */
    int32_t clamp_axis;
    double clamp_value;
    /* traction on boundary faces whose three corners lie on x[load_axis] == load_value */
    int32_t load_axis;
    double load_value;
    int32_t traction_kind;
    double traction[3];     /* uniform traction */
    double traction_scale;  /* random sigma or torsion magnitude */
    uint64_t traction_seed; /* random tractions */
    int32_t load_case;      /* random tractions: the case index */
    double axis_u, axis_v;  /* torsion axis, in the two axes other than load_axis */
    /* material */
    int32_t material_kind;
    double youngs, poisson;             /* uniform, or the shell */
    double core_youngs, core_poisson;   /* shell/core only */
    double shell_depth;                 /* shell/core only */
    int32_t quadrature_points;          /* 4 or 14 */
    /* a tube about the longest axis when wall_cells > 0: wall_cells cells through each wall of
     * wall_thickness on the two axes across, the rest of their divisions over the hollow */
    int32_t wall_cells;
    double wall_thickness;
} ref_workload;

typedef struct ref_mg_options {
    double neighbour_reach;   /* 3.0 */
    int64_t coarse_dofs;      /* 1500 */
    int64_t max_levels;       /* 12 */
    int64_t smoother_degree;  /* 3 */
    double smoother_range;    /* 30.0 */
    int32_t precision;        /* 0 Double, 1 SingleStorage, 2 Single */
    int32_t reserved;
    int64_t threads;          /* 0: min(logical processors / 2, 8), parallel.rs:43-49 */
} ref_mg_options;

typedef struct ref_result {
    int32_t status;           /* 0 converged, 1 not converged, 2 not positive definite, 3 setup error, 4 model error */
    uint64_t iterations;
    double relative_residual;
    int32_t at_rounding_floor;
    uint64_t fingerprint;     /* FNV-1a of every displacement's bits */
    int64_t total_dofs, free_dofs, nonzeros;
    int64_t levels;
    int64_t coarse_dofs;
    double operator_complexity;
    int64_t memory_bytes;
    double assembly_seconds, setup_seconds, solve_seconds;
} ref_result;

void ref_default_workload(ref_workload *w);
void ref_default_mg_options(ref_mg_options *o);

/* Opaque problem: mesh, model, assembled matrix and load built from a workload. */
typedef struct ref_problem ref_problem;

ref_problem *ref_problem_create(const ref_workload *w);
/* The same with prescribed displacement components (Model::prescribe_displacement,
 * crates/solidfem-elastic/src/study.rs:90-99; the last call for a node and axis wins), eliminated
 * symmetrically once assembled (study.rs:567-593); NULL when a component is refused
 * (study.rs:661-683). A workload with clamp_axis < 0 fixes no node; with load_axis < 0 it loads no
 * face. */
ref_problem *ref_problem_create_prescribed(const ref_workload *w, const uint32_t *nodes, const uint32_t *axes,
                                           const double *values, int64_t count);
void ref_problem_destroy(ref_problem *p);

int64_t ref_problem_node_count(const ref_problem *p);
int64_t ref_problem_vertex_count(const ref_problem *p);
int64_t ref_problem_tet_count(const ref_problem *p);
int64_t ref_problem_free_dofs(const ref_problem *p);
int64_t ref_problem_nonzeros(const ref_problem *p);
int64_t ref_problem_boundary_face_count(const ref_problem *p);
/* copies: nodes (3*node_count), tets (10*tet_count), free map (node_count, UINT32_MAX fixed) */
void ref_problem_nodes(const ref_problem *p, double *out);
void ref_problem_tets(const ref_problem *p, uint32_t *out);
void ref_problem_free_map(const ref_problem *p, uint32_t *out);
/* boundary faces: (tet, face, 6 nodes) as 8 u32 each */
void ref_problem_boundary_faces(const ref_problem *p, uint32_t *out);
/* the faces the load was applied to, in order, as (tet, face) pairs */
int64_t ref_problem_loaded_face_count(const ref_problem *p);
void ref_problem_loaded_faces(const ref_problem *p, uint32_t *out, double *tractions);
/* assembled free stiffness, scalar CSR: row_starts (n+1), columns (nnz), values (nnz) */
void ref_problem_csr(const ref_problem *p, int64_t *row_starts, uint32_t *columns, double *values);
void ref_problem_rhs(const ref_problem *p, double *out);
void ref_problem_rigid_modes(const ref_problem *p, double *out); /* 6 per free dof */
/* element stiffness of element `tet` (900 values, row-major) */
int ref_problem_element_stiffness(const ref_problem *p, int64_t tet, double *ke);
int ref_recover(const ref_problem *p, const double *u, double *point_values, double *nodal_values, int64_t *fault_tet);
int ref_problem_quadrature(const ref_problem *p);
/* The error estimate of a converged solve's displacements u (3 per node): recovery, then
 * ErrorEstimation over every element (error_estimate.rs:121-171). element_errors (one per element
 * done), sums = {error_norm_squared, solution_norm_squared, relative_error}, payload = {a, b,
 * elements done}. Returns 0, or the refusal as the product's status code (45 inverted, 46 material,
 * 53 mixed-material patch, 54 coplanar patch, 55 zero strain energy, 56 non-finite). */
int ref_error_estimate(const ref_problem *p, const double *u, double *element_errors, double *sums, int64_t *payload);
/* peak_with_holder over the nodes, then the points, the masks keep (NULL: all); 1 with the peak,
 * 0 when none is inside */
int ref_peak(const double *positions, uint64_t nodes, const double *nodal_vm, uint64_t stride, const double *points, uint64_t count,
             const uint8_t *node_inside, const uint8_t *point_inside, int32_t *at_point, uint64_t *index, double *von_mises,
             double *position);
/* MeshComparison::between (comparison.rs:101-141) of a coarse and a fine solve of two problems:
 * nodal von Mises (`stride` apart, one per node), point values (10 per point, element-major) and
 * the masks of the candidates inside the region (NULL: all). values (13): per side (coarse, fine)
 * the peak's von Mises, position (3) and element volume, then the change, the distance and a
 * refusal's value; integers (8): per side the elements inside, at_point and index, then a
 * refusal's a and b. Returns 0 or the refusal as the product's status code. */
int ref_compare(const ref_problem *coarse, const double *coarse_nodal_vm, uint64_t coarse_stride, const double *coarse_points,
                uint64_t coarse_count, const uint8_t *coarse_node_inside, const uint8_t *coarse_point_inside, const ref_problem *fine,
                const double *fine_nodal_vm, uint64_t fine_stride, const double *fine_points, uint64_t fine_count,
                const uint8_t *fine_node_inside, const uint8_t *fine_point_inside, double *values, int64_t *integers);
/* nodal_rotations: 3 per node; 0, or -1 with the failing element */
int ref_nodal_rotations(const ref_problem *p, const double *u, double *out, int64_t *fault_tet);
double ref_problem_assembly_seconds(const ref_problem *p);

/* Multigrid hierarchy built over the problem's matrix. */
typedef struct ref_multigrid ref_multigrid;
ref_multigrid *ref_multigrid_create(const ref_problem *p, const ref_mg_options *o, int32_t *status);
void ref_multigrid_destroy(ref_multigrid *m);
int64_t ref_multigrid_levels(const ref_multigrid *m);
/* The threads the cycle's row products run on, the caller included (Multigrid::threads,
 * cycle.rs:310-316), and the default count. */
int64_t ref_multigrid_threads(const ref_multigrid *m);
int64_t ref_default_threads(void);
double ref_multigrid_setup_seconds(const ref_multigrid *m);
/* level info: rows, row_block, nnzb, prolongator nnzb (0 on coarsest), has_factor */
void ref_multigrid_level_info(const ref_multigrid *m, int64_t level, int64_t *info5);
/* the level's f64 matrix (as built by setup, before precision rounding) */
void ref_multigrid_level_matrix(const ref_multigrid *m, int64_t level, int64_t *starts, uint32_t *columns, double *values);
void ref_multigrid_level_prolongator(const ref_multigrid *m, int64_t level, int64_t *starts, uint32_t *columns, double *values);
void ref_multigrid_level_inverse_diagonal(const ref_multigrid *m, int64_t level, double *out);
void ref_multigrid_level_spectrum(const ref_multigrid *m, int64_t level, double *lambda_min_max);
void ref_multigrid_level_aggregates(const ref_multigrid *m, int64_t level, uint32_t *out);
/* coarsest factor: n*n row-major lower factor (n = coarsest dofs), 0 if none */
int ref_multigrid_coarse_factor(const ref_multigrid *m, double *out);
/* one V-cycle z = M^-1 r in the hierarchy's precision */
void ref_multigrid_apply(const ref_multigrid *m, const double *r, double *z);
void ref_multigrid_stats(const ref_multigrid *m, int64_t *levels, int64_t *coarse_dofs, double *operator_complexity, int64_t *memory_bytes);

/* Full solve: assembly, multigrid setup, PCG to `tolerance`. displacements: 3*node_count. */
/* The conjugate gradient of the problem's matrix, preconditioned by `mg`, for another right-hand
 * side (free unknowns in, free unknowns out). */
int ref_solve_rhs(const ref_problem *p, const ref_multigrid *mg, const double *rhs, double tolerance, double *x,
                  ref_result *result);
int ref_solve(const ref_workload *w, const ref_mg_options *o, double tolerance,
              double *displacements, ref_result *result);
/* Solve an existing problem + hierarchy; max_iterations 0 = the default cap.
 * iteration_limit > 0 stops after that many iterations (a bounded timing sample). */
int ref_solve_with(const ref_problem *p, const ref_multigrid *m, double tolerance,
                   uint64_t iteration_limit, double *displacements, ref_result *result);

/* Dense lower-level pieces, for kernel-by-kernel parity tests. */
/* Chebyshev smoother on level `level` from a zero guess: x = poly(D^-1 A) D^-1 b (hierarchy precision) */
void ref_multigrid_smooth(const ref_multigrid *m, int64_t level, const double *b, double *x);

/* CsrMatrix::multiply and multiply_rounding_bound (csr.rs:112-154) on plain arrays. */
void ref_csr_multiply(int64_t n, const int64_t *row_starts, const uint32_t *columns, const double *values,
                      const double *x, double *y);
double ref_csr_rounding_bound(int64_t n, const int64_t *row_starts, const uint32_t *columns, const double *values,
                              const double *x);

/* FNV-1a over the little-endian bits of `count` doubles */
uint64_t ref_fingerprint(const double *values, int64_t count);

/* PCG with Jacobi on an arbitrary CSR matrix (the 1-D Laplacian closed form and friends). */
int ref_pcg_jacobi(int64_t n, const int64_t *row_starts, const uint32_t *columns, const double *values,
                   const double *rhs, double tolerance, uint64_t slice, double *x, ref_result *result);

#ifdef __cplusplus
}
#endif
#endif
