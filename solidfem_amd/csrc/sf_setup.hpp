// The multigrid hierarchy's construction on the host: a resumable, cancellable phase machine
// (crates/solidfem-solve/src/multigrid/setup.rs:81-695) whose row-independent phases run on the
// host thread pool in row chunks. Each output row depends only on its own inputs and is written
// in row order, so neither the thread count nor the slicing changes a bit.
#pragma once

#include <array>
#include <chrono>
#include <future>
#include <memory>

#include "sf_host.hpp"

namespace sf {

// The values of a matrix kept elsewhere (a study's assembled system on the device), fetched
// whole or in part when the host needs them.
class ValueSource {
   public:
    virtual ~ValueSource() = default;
    // Every value: row-major blocks in storage order.
    virtual void values(double *out) const = 0;
    // The blocks at the given storage slots, in that order.
    virtual void blocks(const std::vector<int64_t> &slots, double *out) const = 0;
    // Per stored block, the sum of its squared values in storage order, from zero.
    virtual void block_squares(double *out) const = 0;
    // The values on the device, when they are there (on the device of the setup's context).
    virtual const double *device() const { return nullptr; }
};

struct HostLevel {
    // The level's matrix. The finest level of a study leaves its values with the study's system
    // on the device (`source`), and a prolongator built on the device keeps its values there
    // (`prolongator_source`); fetch_values() brings both to the host on first use.
    mutable HostBsr<double> a;
    std::shared_ptr<const ValueSource> source;
    void fetch_values() const;
    std::vector<double> inverse_diagonal;
    double lambda_min = 0, lambda_max = 0;
    bool has_prolongator = false;
    mutable HostBsr<double> prolongator;  // fine rows x coarse blocks (3x6 or 6x6)
    std::shared_ptr<const ValueSource> prolongator_source;
    bool has_factor = false;
    std::vector<double> factor;  // n x n, lower triangle is the Cholesky factor
    std::vector<u32> aggregates;
};

struct SetupOptions {
    double reach = 3.0;
    size_t coarse_dofs = 1500;
    size_t max_levels = 12;
    size_t degree = 3;
    double range = 30.0;
    int precision = SF_PRECISION_SINGLE_STORAGE;
    int coarse_solve = SF_COARSE_REFERENCE;
    int setup_device = SF_SETUP_AUTO;
};

// Runs the numeric phases of a level (the power iteration, the smoothed prolongator, A P and
// P^T A P) somewhere other than the host thread pool, with the host phases' bits. A phase that
// returns false did not run (its rows overflow the accelerator's limits) and runs on the host.
class SetupAccelerator {
   public:
    virtual ~SetupAccelerator() = default;
    virtual bool takes(int level, const HostBsr<double> &a) const = 0;
    // `device_values`, when given, are a's values already on the device; else a's host values.
    virtual void begin(int level, const HostBsr<double> &a, const std::vector<double> &inverse_diagonal, const double *device_values) = 0;
    // `power` is the normalised start vector; returns the last iteration's norm.
    virtual double spectral(std::vector<double> &power, int iterations) = 0;
    virtual bool smooth(const std::vector<u32> &aggregate, u32 count, const std::vector<int64_t> &ns, const std::vector<u32> &nc,
                        const std::vector<double> &tentative, double omega, HostBsr<double> &p) = 0;
    virtual bool product(u32 count) = 0;
    virtual bool galerkin(u32 count, const std::vector<int64_t> &pcol_starts, const std::vector<std::pair<u32, int64_t>> &entries,
                          HostBsr<double> &coarse) = 0;
    // The prolongator's values after a successful smooth(), which leaves only its pattern on the
    // host; the source keeps them after end_level().
    virtual std::shared_ptr<const ValueSource> prolongator_values() = 0;
    // A P from the last successful product(), which keeps it on the device, for a Galerkin
    // product that falls back to the host.
    virtual void product_values(HostBsr<double> &ap) = 0;
    // The coarse matrix's values after a successful galerkin(), which leaves only its pattern
    // on the host (with the unit diagonal where no aggregate direction reaches already applied).
    virtual std::shared_ptr<const ValueSource> coarse_values() = 0;
    virtual void end_level() = 0;
    // The seed pass of aggregation (aggregate.rs:119-136) over the strong neighbour lists, with
    // the serial pass's result: aggregate numbers for the seeds' rows, their seeded flags and the
    // count. False leaves it to the host. It may run while spectral() runs on another thread.
    virtual bool seed(int level, const std::vector<int64_t> &starts, const std::vector<u32> &columns, std::vector<u32> &aggregate,
                      std::vector<char> &seeded, u32 &count) {
        (void)level, (void)starts, (void)columns, (void)aggregate, (void)seeded, (void)count;
        return false;
    }
    // The coarsest level's dense Cholesky factor (block.rs:327-348) of the n x n row-major matrix
    // `a`, in place, with the host loop's result to the bit. False leaves it to the host; a matrix
    // that is not positive definite throws the host's error. It may run while spectral() runs.
    virtual bool factor(int level, size_t n, double *a) {
        (void)level, (void)n, (void)a;
        return false;
    }
};

class MultigridSetup {
   public:
    MultigridSetup(HostBsr<double> matrix, const std::vector<double> &near_null, const std::vector<Vec3> &positions,
                   const SetupOptions &options);
    // Hands the numeric phases of the levels it takes to `accelerator`.
    void attach(std::unique_ptr<SetupAccelerator> accelerator) { accelerator_ = std::move(accelerator); }
    // The finest matrix was given without values: `source` holds them. The host then reads only
    // the diagonal blocks and the block norms unless a phase falls back to the host.
    void hold_values_elsewhere(std::shared_ptr<const ValueSource> source) { source_ = std::move(source); }
    // Up to `budget` block rows of work; true once finished. Throws Error(SF_ERR_MG_*).
    bool step(size_t budget, const Cancel &cancel);
    bool finished() const { return phase_ == Phase::Done; }
    std::vector<HostLevel> take_levels();
    const std::vector<HostLevel> &levels() const { return levels_; }
    const SetupOptions &options() const { return options_; }

   private:
    enum class Phase { Diagonal, Strength, Seed, Join, Sweep, Group, Tentative, Spectral, Filter, Smooth, Product, Galerkin, Factor, Done };
    // Runs rows [cursor, rows) of a phase in chunks, charging `cost` per row. Returns false when
    // paused (budget spent) and throws on cancel.
    template <typename F>
    bool run_rows(size_t rows, size_t &budget, const Cancel &cancel, bool parallel, F &&chunk);
    void begin_level();
    void finish_level();
    void push_final_level();
    void spectral_iteration();
    // The normalised start vector (block.rs:379-382), and the host iteration's work vectors.
    void init_power(bool host_iteration);
    // Every block's Frobenius norm, for the join, when the values are held elsewhere.
    std::vector<double> held_block_norms() const;
    // Brings the current level's values, or its prolongator's, to the host when they are held
    // elsewhere.
    void ensure_values();
    void ensure_prolongator_values();
    // Ends the accelerator's part in the current level, if it has one.
    void release_accelerator_level();
    // SOLIDFEM_SETUP_LOG: wall time per level and phase, printed when the setup finishes.
    void charge_phase(int level);

    SetupOptions options_;
    double reach_ = 3.0;
    size_t retries_ = 0;
    std::vector<Vec3> positions_, coarse_positions_;
    std::vector<HostLevel> levels_;
    HostBsr<double> current_;
    std::vector<double> near_null_, coarse_null_, inverse_diagonal_;
    std::vector<int64_t> strong_starts_;
    std::vector<u32> strong_columns_;
    std::vector<double> closest_;
    std::vector<int64_t> filter_starts_;
    std::vector<u32> filter_columns_;
    std::vector<u32> aggregate_;
    std::vector<char> seeded_;
    u32 aggregate_count_ = 0;
    std::vector<int64_t> member_starts_;
    std::vector<u32> members_;
    std::vector<double> tentative_values_;  // one rb x 6 block per row, in row order
    std::vector<double> power_, scratch_, scratch2_;
    size_t power_done_ = 0;
    double lambda_max_ = 0;
    HostBsr<double> prolongator_, a_times_p_, coarse_;
    std::vector<int64_t> pcol_starts_;
    std::vector<std::pair<u32, int64_t>> pcol_entries_;
    size_t cursor_ = 0;
    std::unique_ptr<SetupAccelerator> accelerator_;
    std::shared_ptr<const ValueSource> source_;  // the current level's values, when not on the host
    std::shared_ptr<const ValueSource> p_source_;  // its prolongator's values, when not on the host
    std::shared_ptr<const ValueSource> c_source_;  // the coarse matrix's values, when not on the host
    std::vector<double> diagonal_blocks_;        // fetched from source_ for the block inverse
    std::vector<double> block_norms_;            // fetched from source_ for the join
    bool on_device_ = false;  // this level's numeric phases run on the accelerator
    bool device_p_ = false;   // the accelerator holds this level's prolongator
    bool device_ap_ = false;  // and its A P
    bool is_coarsest_ = false;
    Phase phase_ = Phase::Diagonal;
    Phase timed_phase_ = Phase::Diagonal;
    int timed_level_ = 0;
    std::chrono::steady_clock::time_point timed_since_;
    std::vector<std::array<double, 14>> phase_seconds_;
    bool cancelled_ = false;
    bool failed_ = false;
    Error failure_{SF_OK, ""};
    // The power iteration of a level the accelerator takes, started once the block inverses are
    // known and run beside the aggregation; the Spectral phase collects it. Declared last, so
    // that it is waited for before anything it reads goes.
    std::future<double> spectral_job_;
    std::future<std::vector<double>> norms_job_;  // the join's block norms, fetched early
};

bool invert_spd(const double *input, size_t n, double *output, double &pivot);

}  // namespace sf
