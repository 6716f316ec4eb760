// Device objects of the solve path: sliced block matrices, the multigrid hierarchy and its
// V-cycle, the conjugate gradient solve and the stiffness assembly, all resident in HBM.
#pragma once

#include <cmath>
#include <functional>
#include <string>

#include "sf_host.hpp"
#include "sf_kernels.hpp"
#include "sf_setup.hpp"

namespace sf {

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    explicit Context(int dev);
    ~Context();
};

// The row order and slot layout of a sliced matrix, planned on the host.
struct SlicedPlan {
    int rb = 3, cb = 3, npw = 21;
    size_t rows = 0, blocks = 0;
    uint32_t slices = 0;
    std::vector<uint32_t> slice_start, row_len, perm, cols, block_of_slot;
};
SlicedPlan plan_sliced(const int64_t *starts, const u32 *columns, size_t rows, int rb, int cb);

template <typename T>
struct DeviceSliced {
    int rb = 3, cb = 3;
    size_t rows = 0, block_cols = 0, blocks = 0;
    uint32_t slices = 0;
    DeviceBuffer<uint32_t> slice_start, row_len, perm, cols;
    DeviceBuffer<T> vals;
    int batch_spmv = 4;  // blocks in flight per lane, chosen by tune_batch
    int batch_cheb = 4;
    mutable int batch_mm[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // batched product, per log2 of the column group
    mutable uint32_t passes = 0;  // launches so far, for the alternating traversal
    // The view one launch takes; each call is one pass (see dev::slice_of for the orders).
    dev::Sliced<T> view() const {
        uint32_t order = xcd_order();
        if (ping_pong() && (passes++ & 1u)) order |= dev::ORDER_REVERSE;
        return {slice_start.ptr, row_len.ptr, perm.ptr, cols.ptr, vals.ptr, slices, order};
    }
    // The view the batched (several-column) kernels take: with SOLIDFEM_GROUP_RUNS=1 their
    // workgroups run in contiguous runs per XCD.
    dev::Sliced<T> batch_view() const {
        dev::Sliced<T> v = view();
        const int g = group_runs();
        if (g == 1) v.order |= dev::ORDER_GROUP_RUNS;
        else if (g > 1) v.order |= dev::ORDER_GROUP_CHUNKS | ((uint32_t)__builtin_ctz((unsigned)g) << dev::ORDER_CHUNK_SHIFT);
        return v;
    }
    static int group_runs();
    static uint32_t xcd_order();  // SOLIDFEM_XCD_ORDER: 0 none, 1 contiguous runs, 2 whole windows
    static bool ping_pong();      // SOLIDFEM_PING_PONG: alternate passes walk the slices backwards
    size_t dofs() const { return rows * rb; }
    // Bytes one product reads by the algorithmic count: values, block columns, row offsets.
    size_t algorithmic_matrix_bytes() const { return blocks * (rb * cb * sizeof(T) + 4) + (rows + 1) * 4; }
};

// A block matrix in plain block rows (dev::Bsr) in precision T.
template <typename T>
struct DeviceBsr {
    int rb = 0, cb = 0;
    uint32_t rows = 0;
    DeviceBuffer<uint32_t> starts, cols;
    DeviceBuffer<T> vals;
    mutable uint32_t passes = 0;  // launches so far, for the alternating traversal
    // Every other pass takes each XCD's rows in reverse order (SOLIDFEM_ROWS_PING_PONG=0: never).
    dev::Bsr<T> view() const {
        const uint32_t reverse = rows_ping_pong() ? (passes++ & 1u) : 0u;
        return {starts.ptr, cols.ptr, vals.ptr, rows, reverse};
    }
    static bool rows_ping_pong();
};
template <typename T>
void upload_bsr(DeviceBsr<T> &out, const HostBsr<double> &m, hipStream_t s);
// The same from `pattern` with its values already on the device (f64, row-major blocks).
template <typename T>
void build_bsr(DeviceBsr<T> &out, const HostBsr<double> &pattern, const double *values, hipStream_t s);
// y = A x (epilogue per dev::Epilogue) with `rows_per_wave` block rows per wavefront
// (long_rows_kernel: 1, 2 or 4).
template <typename S, typename V>
void launch_long_rows(const DeviceBsr<S> &a, const V *x, V *y, const V *b, int epi, hipStream_t s,
                      dev::Start<S, V> start = {nullptr, (V)0, nullptr, nullptr}, int rows_per_wave = 1);

// Builds a sliced matrix from row-major f64 blocks already on the device (`src_values`).
template <typename T>
void build_sliced(DeviceSliced<T> &out, const SlicedPlan &plan, size_t block_cols, const double *src_values, hipStream_t s);
// ... or from host blocks.
template <typename T>
void upload_sliced(DeviceSliced<T> &out, const HostBsr<double> &m, hipStream_t s);

// y = A x (epilogue per dev::Epilogue) for any supported block shape, with the matrix's batch.
template <typename S, typename V>
void launch_spmv(const DeviceSliced<S> &a, const V *x, V *y, const V *b, int epi, hipStream_t s, int batch = 0,
                 dev::Start<S, V> start = {nullptr, (V)0, nullptr, nullptr});

// The batch depth (blocks in flight per lane, one of `candidates`) whose launches
// `launch(batch)` run fastest; every depth gives the same bits. SOLIDFEM_BATCH forces one.
// The tuning variant forced on every tuning that has it among its candidates (0: none).
std::atomic<int> &forced_variant();
int tune_batch(hipStream_t s, const std::function<void(int)> &launch, const std::vector<int> &candidates = {1, 2, 4, 8},
               const std::string &what = "", const std::string &memo = "");
// The candidate whose launches `launch(candidate)` run fastest (timed with HIP events on `s`).
// With a `memo` key (a kernel and a matrix shape class), the first choice made for that key and
// candidate list is kept for the process and returned again without timing: matrices of one
// class (a new system in a server, a study rebuilt) skip the timing. Every choice gives the
// same bits.
int fastest(hipStream_t s, const std::function<void(int)> &launch, const std::vector<int> &candidates,
            const std::string &what = "", const std::string &memo = "");
// A matrix's shape class for the tuning memo: block shape, value size, and the binary orders of
// its block rows and of its mean blocks per row.
template <typename T>
std::string shape_class(const DeviceSliced<T> &a) {
    auto order = [](double v) { return v < 1 ? 0 : (int)std::floor(std::log2(v)); };
    return std::to_string(a.rb) + "x" + std::to_string(a.cb) + "/" + std::to_string(sizeof(T)) + "/" +
           std::to_string(order((double)a.rows)) + "/" + std::to_string(order(2.0 * (double)a.blocks / std::max<double>(1.0, (double)a.rows)));
}

// ---- several right-hand sides ------------------------------------------------------------------

// Columns one batched solve carries: a group of lanes per block row, one lane per column.
constexpr uint32_t MAX_COLUMNS = 64;
// A batched kernel's variant code: below NODE_FORM, the lane-group form with that depth (blocks in
// flight); from NODE_FORM, the one-wavefront-per-block-row form with depth (code - NODE_FORM);
// from PAIR_FORM, QUAD_FORM and OCT_FORM, the lane-group form whose lanes take two, four or
// eight consecutive columns each (for M a multiple of that; otherwise the one-column form of the
// same depth).
constexpr int NODE_FORM = 16;
constexpr int PAIR_FORM = 32;
constexpr int QUAD_FORM = 48;
constexpr int OCT_FORM = 64;  // eight columns per lane
constexpr int STAGED_FORM = 80;  // products only: one workgroup per block row, chunks staged in LDS
inline int form_columns(int code) {
    return code >= STAGED_FORM ? 1 : code >= OCT_FORM ? 8 : code >= QUAD_FORM ? 4 : code >= PAIR_FORM ? 2 : 1;
}
inline int form_depth(int code) {
    return code >= STAGED_FORM ? 0 : code >= OCT_FORM ? code - OCT_FORM : code >= QUAD_FORM ? code - QUAD_FORM : code >= PAIR_FORM ? code - PAIR_FORM : code >= NODE_FORM ? code - NODE_FORM : code;
}
inline bool node_form(int code) { return code >= NODE_FORM && code < PAIR_FORM; }
inline int columns_log2(uint32_t m) {
    int g = 0;
    while ((1u << g) < m) ++g;
    return g;
}
// Y = A X for M row-interleaved columns (epilogue per dev::Epilogue), `batch` blocks in flight
// (1, 2 or 4; 0 = the matrix's tuned depth for this column count).
template <typename S, typename V>
void launch_spmm(const DeviceSliced<S> &a, uint32_t M, const V *x, V *y, const V *b, int epi, hipStream_t s, int batch = 0);
// The batch depths worth trying for a matrix's batched product.
// The variant codes a batched product (`product`) or Chebyshev step is timed over for M columns.
std::vector<int> spmm_candidates(int rb, int cb, size_t value_size, uint32_t M, bool product = true, bool staged = false);
// Whether the staged form (STAGED_FORM) is a candidate for products by `a` with M columns of
// `vector_size` bytes: rows of at least 40 blocks on average (or the form forced), at most 256
// (row entry, column) pairs per row, and the staged entries of x within 48 KiB of LDS.
template <typename S>
bool staged_candidate(const DeviceSliced<S> &a, uint32_t M, size_t vector_size);
// Tunes a matrix's product on scratch vectors of its shape.
template <typename S, typename V>
void tune_spmv(DeviceSliced<S> &a, int epi, hipStream_t s);

// ---- launch timing ------------------------------------------------------------------------------

// Brackets launches of a chosen kernel with HIP events on the launching stream and accumulates
// their durations and algorithmic bytes (read back on request).
struct Profiler {
    static constexpr int KERNELS = 5;
    struct Pending {
        hipEvent_t start, stop;
        double bytes;
        u64 launches;  // launches of the kernel between the two events (back to back)
    };
    struct Slot {
        bool on = false;
        std::vector<Pending> pending;
        std::vector<hipEvent_t> pool;
        u64 launches = 0;
        double seconds = 0, bytes = 0;
    };
    Slot slots[KERNELS];
    std::mutex lock;
    static Profiler &get();
    // Returns an index into the pending list, or -1 when off.
    int begin(int kernel, hipStream_t s, double bytes, u64 launches = 1);
    void end(int kernel, int token, hipStream_t s);
    void collect(int kernel);
    void reset();
    bool any_on() const {
        for (const Slot &k : slots)
            if (k.on) return true;
        return false;
    }
};

// ---- multigrid --------------------------------------------------------------------------------

class DeviceMultigrid {
   public:
    virtual ~DeviceMultigrid() = default;
    // z = M^-1 r, device pointers, f64 in and out.
    virtual void apply(const double *r, double *z, hipStream_t s) = 0;
    // One smoother application on `level` from a zero guess (f64 device vectors in and out).
    virtual void smooth_level(size_t level, const double *b, double *x, hipStream_t s) = 0;
    virtual void coarse_solve(const double *b, double *x, hipStream_t s) = 0;
    // Z = M^-1 R for M row-interleaved f64 columns (entry (i, m) at i * M + m); column m is bit for
    // bit apply() of column m.
    virtual void apply_batch(const double *r, double *z, uint32_t columns, hipStream_t s) = 0;
    // The tuned batch depth of a level's fused Chebyshev step (kind 0) or residual product (kind 1).
    virtual int tuned_depth(size_t level, int kind) const = 0;
    int precision = SF_PRECISION_SINGLE_STORAGE;
    int coarse_mode = SF_COARSE_REFERENCE;
    size_t degree = 3;
    std::vector<HostLevel> host;  // the f64 hierarchy as set up (for inspection and export)
    sf_multigrid_stats stats{};
    size_t dofs() const { return host.empty() ? 0 : host[0].a.dofs(); }
    // Algorithmic HBM bytes of one V-cycle (every array each kernel touches, counted once).
    double cycle_bytes = 0;
    // apply() replays a HIP graph of the cycle's launches, captured once per (r, z) pair, instead
    // of launching them one by one (not while kernel timing is on). The same kernels, the same bits.
    bool use_graph = false;
};
// `finest`, when given, is level 0's matrix already on the device in f64 (a study's system
// matrix, the same pattern and values): the hierarchy's finest level is converted from it instead
// of being planned and uploaded from the host copy again.
std::unique_ptr<DeviceMultigrid> make_device_multigrid(std::vector<HostLevel> levels, size_t degree, int precision,
                                                       int coarse_mode, hipStream_t s, const DeviceSliced<double> *finest = nullptr);

// ---- system matrix and conjugate gradient ---------------------------------------------------

// The system matrix on the device: 3x3 blocks when the matrix groups into nodal blocks, else
// scalar rows. Keeps the host scalar-CSR view for export.
struct DeviceSystem {
    size_t size = 0;  // scalar rows
    int rb = 3;
    HostBsr<double> host;  // block (or scalar) rows with f64 values, as assembled
    DeviceBuffer<double> values;  // row-major blocks on the device (assembly output)
    DeviceSliced<double> sliced;
    void upload_from_host(hipStream_t s);
    void build_from_device_values(hipStream_t s);
};

// The Jacobi preconditioner's inverse diagonal of a system (Jacobi::new,
// crates/solidfem-solve/src/preconditioner.rs:53-66); throws SF_ERR_JACOBI_DIAGONAL.
std::vector<double> jacobi_inverse(const std::shared_ptr<Context> &ctx, const DeviceSystem &a);

enum class PcgStatus { Active, Converged, Cancelled, NotPositiveDefinite, NotConverged };

class DevicePcg {
   public:
    DevicePcg(std::shared_ptr<Context> ctx, const DeviceSystem &a, DeviceMultigrid *mg, const std::vector<double> &rhs,
              const sf_pcg_options &options, const sf_preconditioner *user = nullptr);
    // Pcg::new again on the buffers already held, optionally with a new right-hand side.
    void restart(const double *new_rhs);
    ~DevicePcg();
    // Pcg::step: up to `max` iterations, cancel polled before each. Throws Error on failure.
    int step(u64 max, const Cancel &cancel);
    void estimate(double *out) const;
    sf_progress progress() const;
    bool converged() const { return status_ == PcgStatus::Converged; }
    size_t size() const { return n_; }
    const double *x_device() const { return x_.ptr; }
    u64 iterations() const { return iterations_; }

   private:
    void precondition(const double *r, double *z);
    void dot(const double *a, const double *b, double *result, int final_kind = dev::FINAL_STORE, dev::PcgScalars *mirror = nullptr,
             uint32_t tag = 0);
    void wait_for_tag(uint32_t tag);
    void sum(const double *a, double *result);
    void iterate();
    void reset_direction();
    void initialise();
    dev::PcgScalars read_scalars();

    std::shared_ptr<Context> ctx_;  // kept alive while this solve may launch on its stream
    const DeviceSystem &a_;
    DeviceMultigrid *mg_;
    size_t n_;
    int dot_order_;
    double tolerance_, rhs_norm_ = 0, relative_ = 0;
    u64 iterations_ = 0, max_iterations_;
    bool floor_ = false;
    PcgStatus status_ = PcgStatus::Active;
    Error failure_{SF_OK, ""};
    DeviceBuffer<double> b_, x_, r_, z_, p_, ap_, tmp_, tmp2_, jacobi_;
    DeviceBuffer<dev::PcgScalars> scalars_;
    DeviceBuffer<double> partials_, partials2_, norm_;
    uint32_t rz_slot_ = 0;  // TREE order: the PcgScalars slot holding the current r . z (rz or rz_next)
    dev::PcgScalars *pinned_ = nullptr;  // fine-grained host memory the r.r kernel copies the scalars into
    dev::PcgScalars *mirror_ = nullptr;  // its device address
    uint32_t readbacks_ = 0;             // tags of the copies so far
    sf_preconditioner user_{};           // a caller's preconditioner (apply null: none)
    double *host_r_ = nullptr, *host_z_ = nullptr;  // pinned staging of a host-side preconditioner
};

// Several load cases over one matrix and one preconditioner: every column runs the recurrence of
// Pcg (crates/solidfem-solve/src/pcg.rs:113-283) on its own scalars, in lockstep with the others
// so that each matrix pass serves all columns. A column's operations are those of DevicePcg in
// the same order, so its iterates are bit for bit those of solving it alone; a column that
// converges, fails or is cancelled is frozen while the rest go on.
class DevicePcgBatch {
   public:
    DevicePcgBatch(std::shared_ptr<Context> ctx, const DeviceSystem &a, DeviceMultigrid *mg, const double *rhs_columns, uint32_t columns,
                   const sf_pcg_options &options);
    ~DevicePcgBatch();
    // Pcg::new again on the held buffers, optionally with new right-hand sides (column-major).
    void restart(const double *rhs_columns);
    // Up to `max` lockstep iterations; cancel is polled before each. Returns the running columns.
    uint32_t step(u64 max, const Cancel &cancel);
    void estimate(double *columns_out) const;  // column-major
    void estimate_column(uint32_t column, double *out) const;
    // Pcg::new for the listed columns with new right-hand sides (host, n each), between steps;
    // the other columns go on unchanged. A column of zeros converges at once and stays idle.
    void replace_columns(const std::vector<uint32_t> &columns, const std::vector<const double *> &rhs);
    void progress(sf_column_progress *out) const;
    size_t device_bytes() const;
    uint32_t columns() const { return m_; }
    size_t size() const { return n_; }

   private:
    struct Column {
        u64 iterations = 0;
        double rhs_norm = 0, relative = 0;
        bool floor = false;
        PcgStatus status = PcgStatus::Active;
        Error failure{SF_OK, ""};
    };
    void initialise();
    void iterate();
    void precondition(const double *r, double *z);
    void dot(const double *a, const double *b, size_t field, int kind, const uint32_t *mask, int mode = 0);
    void reset_direction(const uint32_t *mask);
    void upload_flags(uint32_t *device, const std::vector<uint32_t> &flags);
    const dev::PcgScalars *read_scalars();
    std::vector<uint32_t> active_flags() const;

    std::shared_ptr<Context> ctx_;  // kept alive while this solve may launch on its stream
    const DeviceSystem &a_;
    DeviceMultigrid *mg_;
    size_t n_;
    uint32_t m_;
    int dot_order_;
    double tolerance_;
    u64 max_iterations_;
    std::vector<Column> col_;
    DeviceBuffer<double> b_, x_, r_, z_, p_, ap_, tmp_, tmp2_, jacobi_, stage_;
    DeviceBuffer<dev::PcgScalars> scalars_;
    DeviceBuffer<double> partials_;
    DeviceBuffer<uint32_t> active_, mask_;
    dev::PcgScalars *pinned_ = nullptr;
    uint32_t *pinned_flags_ = nullptr;
    hipEvent_t copied_ = nullptr;
};

// ---- stress recovery --------------------------------------------------------------------------

// A resumable stress recovery (Recovery, crates/solidfem-elastic/src/results.rs:103-124 and
// :307-576, with element.rs:135-192 and patch.rs) of one snapshot of nodal displacements (3 per
// mesh node), on the device: every element in order, then one patch fit per mesh vertex in order.
// Slices of any size give the bits of one run: an element's values depend on that element alone,
// a fit on its vertex's patch once every element is done, and a node's worst incident value is
// taken over its elements in element order, once, before the first fit. The field: per
// (element, point) 10 values -- position 3, stress [xx, yy, zz, xy, yz, zx], von Mises -- and per
// mesh node 7 -- the recovered stress and its von Mises.
class DeviceRecovery {
   public:
    DeviceRecovery(const Model &model, const double *u, bool converged, hipStream_t s);
    // Up to `max_work` more units (an element, then a vertex's patch fit); `cancel` (may be empty)
    // is read before each. Returns whether every unit is done. A cancel throws
    // SF_ERR_MODEL_CANCELLED with the units before it done; a faulting element throws with its
    // index, the elements before it done and it the next to do.
    bool advance(size_t max_work, const std::function<bool()> &cancel);
    bool finished() const { return next_ == elements_ && next_vertex_ == vertices_; }
    size_t elements_done() const { return next_; }
    size_t fits_done() const { return next_vertex_; }
    size_t elements() const { return elements_; }
    size_t vertices() const { return vertices_; }
    size_t nodes() const { return nodes_; }
    int points_per_element() const { return nq_; }
    bool converged() const { return converged_; }
    const std::vector<double> &displacements() const { return u_; }
    // The finished field (either output may be null); SF_ERR_MODEL_NOT_FINISHED before then.
    void field(double *point_values, double *nodal_values);
    // On the device: the displacements (3 per node) and the patch fits (29 per vertex: a flag --
    // 1 fitted, 0 flat, -1 more than one material -- the centre, the scale and 4 x 6 coefficients).
    const double *device_displacements() const { return u_d_.ptr; }
    const double *device_fits() const { return fit_.ptr; }

   private:
    void run_elements(size_t count);
    void run_fits(size_t count);
    void ensure_worst();

    hipStream_t s_;
    size_t elements_, nodes_, vertices_;
    int nq_;
    MaterialField field_;
    bool converged_;
    std::vector<double> u_;
    size_t next_ = 0, next_vertex_ = 0;
    bool worst_done_ = false, nodal_done_ = false;
    DeviceBuffer<double> xyz_, u_d_, points_, material_, nodevals_, worst_, fit_, nodal_;
    DeviceBuffer<double> field_table_;  // a caller's field at the rule's points
    DeviceBuffer<uint32_t> tets_, nstart_, nat_, vstart_, vat_, ends_;
    DeviceBuffer<int8_t> nearest_;
    DeviceBuffer<int32_t> fault_;
};

// ErrorEstimation (crates/solidfem-elastic/src/error_estimate.rs:80-238) of a recovery it borrows
// (which must outlive it), over `model`'s mesh and material: per element, the 14-point integrals of
// the recovered minus the element stress and of the element stress in the compliance norm on the
// device; the running sums in element order on the host, so any slicing gives one run's bits.
class DeviceErrorEstimation {
   public:
    DeviceErrorEstimation(const Model &model, DeviceRecovery &recovery, hipStream_t s);
    // Up to `max_elements` more elements, `cancel` (may be empty) read before each; whether every
    // element is done. Refuses a recovery of another mesh (SF_ERR_MODEL_RECOVERY_MESH_MISMATCH),
    // an unfinished or preview one (SF_ERR_MODEL_NOT_FINISHED); a cancel throws
    // SF_ERR_MODEL_CANCELLED and an element's refusal its error, the elements before it done.
    bool advance(size_t max_elements, const std::function<bool()> &cancel);
    size_t elements_done() const { return next_; }
    size_t elements() const { return elements_; }
    // ErrorEstimation::estimate: element errors (one per element; may be null) and the sums.
    void estimate(double *element_errors, double *error_norm_squared, double *solution_norm_squared,
                  double *relative_error) const;

   private:
    bool matches() const;
    void run(size_t count);
    Error refusal(unsigned long long fault, size_t first) const;

    DeviceRecovery &recovery_;
    hipStream_t s_;
    size_t elements_, nodes_, vertices_;
    MaterialField field_;
    DeviceBuffer<double> xyz_, recovered_, terms_, sums_;
    DeviceBuffer<double> field_table_;  // a caller's field at the 14-point rule's points
    DeviceBuffer<uint32_t> tets_;
    DeviceBuffer<unsigned long long> fault_;
    size_t next_ = 0;
    std::vector<double> element_errors_;
    double error_sum_ = 0.0, energy_sum_ = 0.0;
};

// The whole recovery in one call; either output may be null.
void recover_stress(const Model &model, const double *u, double *point_values, double *nodal_values, hipStream_t s);

// nodal_rotations (results.rs:262-305): half the curl of the displacement at every mesh node, the
// element values at the node weighted by their Jacobian determinants and summed in element order.
void nodal_rotations(const Model &model, const double *u, double *rotations, hipStream_t s);

// ---- assembly -------------------------------------------------------------------------------

// Element stiffness (crates/solidfem-elastic/src/element.rs:98-133) on the device, and the
// gather of each block's contributions in ascending element order (assembly.rs:87-130).
struct DeviceAssembly {
    DeviceBuffer<double> nodes;     // 3 per mesh node
    DeviceBuffer<uint32_t> tets;    // 10 per element
    DeviceBuffer<int64_t> contribution_starts;
    DeviceBuffer<uint32_t> contributions;
    DeviceBuffer<double> ke;        // 55 upper blocks x 9 per element of a chunk
    DeviceBuffer<int32_t> fault;    // first failing element, point and kind
    DeviceBuffer<double> field_table;  // a caller's field at the rule's points
    size_t elements = 0;
    MaterialField field;
    int quadrature = 4;
    void prepare(const Model &model, const AssemblyPattern &pattern, hipStream_t s);
    // Adds elements [first, last) into `values` (row-major 3x3 blocks, zeroed before the first).
    void add_elements(size_t first, size_t last, double *values, size_t blocks, hipStream_t s);
    // The 900-value element stiffness of elements [first, first + count) into host `ke900`.
    void element_matrices(size_t first, size_t count, double *ke900, hipStream_t s);
    void check_fault(hipStream_t s);
};

}  // namespace sf
