// The multigrid hierarchy's construction on the host: a resumable, cancellable phase machine
// (crates/solidfem-solve/src/multigrid/setup.rs:81-695) whose row-independent phases run on the
// host thread pool in row chunks. Each output row depends only on its own inputs and is written
// in row order, so neither the thread count nor the slicing changes a bit.
#pragma once

#include "sf_host.hpp"

namespace sf {

struct HostLevel {
    HostBsr<double> a;
    std::vector<double> inverse_diagonal;
    double lambda_min = 0, lambda_max = 0;
    bool has_prolongator = false;
    HostBsr<double> prolongator;  // fine rows x coarse blocks (3x6 or 6x6)
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
};

class MultigridSetup {
   public:
    MultigridSetup(HostBsr<double> matrix, const std::vector<double> &near_null, const std::vector<Vec3> &positions,
                   const SetupOptions &options);
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
    bool is_coarsest_ = false;
    Phase phase_ = Phase::Diagonal;
    bool cancelled_ = false;
    bool failed_ = false;
    Error failure_{SF_OK, ""};
};

bool invert_spd(const double *input, size_t n, double *output, double &pivot);

}  // namespace sf
