// Host multigrid setup: the phases of crates/solidfem-solve/src/multigrid/setup.rs,
// setup_aggregation.rs, aggregate.rs and transfer.rs, in the reference's floating-point order,
// with the row-independent phases on the host thread pool.
#include "sf_setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>

namespace sf {

namespace {

constexpr size_t MODES = 6;                 // multigrid/mod.rs:124
constexpr size_t DENSE_LIMIT = 3000;        // setup.rs:28
constexpr double MAX_COARSE_FRACTION = 0.7; // setup.rs:32
constexpr double RETRY_WIDENING = 1.5;      // setup.rs:36
constexpr size_t MAX_RETRIES = 3;           // setup.rs:37
constexpr size_t POWER_ITERATIONS = 20;     // setup.rs:41
constexpr double SPECTRAL_MARGIN = 1.15;    // setup.rs:45
constexpr double RANK_TOLERANCE = 1e-10;    // transfer.rs:12
constexpr u32 NONE = 0xffffffffu;
constexpr size_t CHUNK = 4096;              // rows between cancel polls in the serial phases
constexpr size_t PARALLEL_CHUNK = 32768;    // and in the row-parallel ones (fewer pool dispatches)

// C += A B for row-major dense blocks (block.rs:229-245).
inline void gemm_acc(double *c, const double *a, size_t m, size_t k, const double *b, size_t n) {
    for (size_t r = 0; r < m; ++r)
        for (size_t mid = 0; mid < k; ++mid) {
            const double s = a[r * k + mid];
            for (size_t col = 0; col < n; ++col) c[r * n + col] += s * b[mid * n + col];
        }
}

// A sparse row gathered from columns in any order: blocks summed in arrival order per column,
// read back by ascending column (block.rs:249-302).
struct RowAccumulator {
    std::vector<u32> slot, touched;
    std::vector<double> data;
    size_t len;
    RowAccumulator(size_t columns, size_t block_len) : slot(columns, NONE), len(block_len) {}
    void clear() {
        for (u32 c : touched) slot[c] = NONE;
        touched.clear();
        data.clear();
    }
    double *at(u32 column) {
        u32 s = slot[column];
        if (s == NONE) {
            s = (u32)touched.size();
            slot[column] = s;
            touched.push_back(column);
            data.resize(data.size() + len, 0.0);
        }
        return &data[(size_t)s * len];
    }
    void sort() { std::sort(touched.begin(), touched.end()); }
    const double *get(u32 column) const { return &data[(size_t)slot[column] * len]; }
};

// Per-row outputs of a parallel chunk, appended to the matrix in row order afterwards.
struct RowOut {
    std::vector<u32> cols;
    std::vector<double> vals;
};

inline double dist2(const std::vector<Vec3> &p, size_t r, size_t c) {
    const double dx = p[c][0] - p[r][0], dy = p[c][1] - p[r][1], dz = p[c][2] - p[r][2];
    return dx * dx + dy * dy + dz * dz;
}

inline double frobenius(const double *b, size_t n) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += b[i] * b[i];
    return std::sqrt(s);
}

template <typename T>
void append_rows(HostBsr<T> &m, std::vector<RowOut> &rows) {
    for (auto &r : rows) {
        m.columns.insert(m.columns.end(), r.cols.begin(), r.cols.end());
        m.values.insert(m.values.end(), r.vals.begin(), r.vals.end());
        m.starts.push_back((int64_t)m.columns.size());
    }
}

// y = A x over rows [first, last) in block order (block.rs:136-154).
void bsr_multiply_rows(const HostBsr<double> &a, const double *x, double *y, size_t first, size_t last) {
    const size_t rb = a.rb, cb = a.cb;
    for (size_t row = first; row < last; ++row) {
        double *out = y + row * rb;
        for (size_t p = 0; p < rb; ++p) out[p] = 0.0;
        for (int64_t at = a.starts[row]; at < a.starts[row + 1]; ++at) {
            const double *b = a.block((size_t)at);
            const double *xc = x + (size_t)a.columns[(size_t)at] * cb;
            for (size_t p = 0; p < rb; ++p) {
                double s = out[p];
                for (size_t q = 0; q < cb; ++q) s += b[p * cb + q] * xc[q];
                out[p] = s;
            }
        }
    }
}

}  // namespace

// SPD block inverse through the inverted Cholesky factor (block.rs:186-224).
bool invert_spd(const double *input, size_t n, double *output, double &pivot) {
    double l[36] = {0}, inv[36] = {0};
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j <= i; ++j) {
            double s = input[i * n + j];
            for (size_t k = 0; k < j; ++k) s -= l[i * n + k] * l[j * n + k];
            if (i == j) {
                if (!std::isfinite(s) || s <= 0.0) {
                    pivot = s;
                    return false;
                }
                l[i * n + j] = std::sqrt(s);
            } else {
                l[i * n + j] = s / l[j * n + j];
            }
        }
    for (size_t c = 0; c < n; ++c)
        for (size_t i = c; i < n; ++i) {
            double s = i == c ? 1.0 : 0.0;
            for (size_t k = c; k < i; ++k) s -= l[i * n + k] * inv[k * n + c];
            inv[i * n + c] = s / l[i * n + i];
        }
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) {
            double s = 0.0;
            for (size_t k = std::max(i, j); k < n; ++k) s += inv[k * n + i] * inv[k * n + j];
            output[i * n + j] = s;
        }
    return true;
}

MultigridSetup::MultigridSetup(HostBsr<double> matrix, const std::vector<double> &near_null,
                               const std::vector<Vec3> &positions, const SetupOptions &options)
    : options_(options), reach_(options.reach), positions_(positions), current_(std::move(matrix)), near_null_(near_null) {
    if (near_null_.size() != MODES * current_.dofs())
        throw Error(SF_ERR_MG_NEAR_NULL_LENGTH, "the rigid-body modes have the wrong length", (int64_t)(MODES * current_.dofs()),
                    (int64_t)near_null_.size());
    if (positions_.size() != current_.rows)
        throw Error(SF_ERR_MG_POSITION_LENGTH, "the block positions have the wrong length", (int64_t)current_.rows,
                    (int64_t)positions_.size());
}

std::vector<HostLevel> MultigridSetup::take_levels() {
    if (phase_ != Phase::Done) throw Error(SF_ERR_MG_NOT_FINISHED, "the multigrid setup has not finished");
    return std::move(levels_);
}

template <typename F>
bool MultigridSetup::run_rows(size_t rows, size_t &budget, const Cancel &cancel, bool parallel, F &&chunk) {
    while (cursor_ < rows) {
        if (budget == 0) return false;
        if (cancel()) {
            cancelled_ = true;
            throw Error(SF_ERR_MG_CANCELLED, "the multigrid setup was cancelled");
        }
        size_t n = std::min({rows - cursor_, budget, parallel ? PARALLEL_CHUNK : CHUNK});
        chunk(cursor_, cursor_ + n, parallel);
        cursor_ += n;
        budget -= n;
    }
    cursor_ = 0;
    return true;
}

void MultigridSetup::spectral_iteration() {
    // setup.rs:409-423: v <- D^-1 A v / ||D^-1 A v||
    const size_t rows = current_.rows, rb = current_.rb;
    parallel_for(rows, 2048, [&](size_t f, size_t l) {
        bsr_multiply_rows(current_, power_.data(), scratch_.data(), f, l);
        for (size_t row = f; row < l; ++row) {
            const double *d = &inverse_diagonal_[row * rb * rb];
            for (size_t p = 0; p < rb; ++p) {
                double s = 0.0;
                for (size_t q = 0; q < rb; ++q) s += d[p * rb + q] * scratch_[row * rb + q];
                scratch2_[row * rb + p] = s;
            }
        }
    });
    double s = 0.0;
    for (double v : scratch2_) s += v * v;
    const double norm = std::sqrt(s);
    lambda_max_ = norm;
    parallel_for(power_.size(), 16384, [&](size_t f, size_t l) {
        for (size_t i = f; i < l; ++i) power_[i] = scratch2_[i] / norm;
    });
}

void MultigridSetup::charge_phase(int level) {
    static const bool log = std::getenv("SOLIDFEM_SETUP_LOG") != nullptr;
    if (!log) return;
    const auto now = std::chrono::steady_clock::now();
    if (phase_seconds_.size() <= (size_t)timed_level_) phase_seconds_.resize((size_t)timed_level_ + 1, {});
    phase_seconds_[(size_t)timed_level_][(size_t)timed_phase_] += std::chrono::duration<double>(now - timed_since_).count();
    timed_since_ = now;
    timed_level_ = level;
    timed_phase_ = phase_;
    if (phase_ != Phase::Done) return;
    static const char *names[] = {"diagonal", "strength", "seed", "join", "sweep", "group", "tentative",
                                  "spectral", "filter", "smooth", "product", "galerkin", "factor"};
    for (size_t l = 0; l < phase_seconds_.size(); ++l) {
        std::fprintf(stderr, "[setup] level %zu:", l);
        for (size_t p = 0; p < 13; ++p)
            if (phase_seconds_[l][p] > 0) std::fprintf(stderr, " %s %.1f ms", names[p], 1e3 * phase_seconds_[l][p]);
        std::fprintf(stderr, "\n");
    }
}

bool MultigridSetup::step(size_t budget, const Cancel &cancel) {
    if (failed_) throw failure_;
    if (cancelled_) throw Error(SF_ERR_MG_CANCELLED, "the multigrid setup was cancelled");
    timed_since_ = std::chrono::steady_clock::now();
    struct Charge {
        MultigridSetup *s;
        ~Charge() { if (s->phase_ != Phase::Done) s->charge_phase((int)s->levels_.size()); }
    } charge_on_exit{this};
    try {
        for (;;) {
            const size_t rows = current_.rows, rb = current_.rb;
            const int level = (int)levels_.size();
            if (phase_ != timed_phase_ || level != timed_level_) charge_phase(level);
            switch (phase_) {
                case Phase::Done:
                    return true;
                case Phase::Diagonal: {
                    // setup.rs:269-321
                    const bool held = source_ && current_.values.empty() && current_.blocks() > 0;
                    if (cursor_ == 0) {
                        inverse_diagonal_.assign(rows * rb * rb, 0.0);
                        if (held) {
                            // only the diagonal blocks come to the host (a missing one reads slot
                            // 0 and is refused below before use)
                            std::vector<int64_t> slots(rows, 0);
                            parallel_for(rows, 4096, [&](size_t a, size_t b) {
                                for (size_t row = a; row < b; ++row) {
                                    auto first = current_.columns.begin() + current_.starts[row];
                                    auto last = current_.columns.begin() + current_.starts[row + 1];
                                    auto it = std::lower_bound(first, last, (u32)row);
                                    if (it != last && *it == (u32)row) slots[row] = it - current_.columns.begin();
                                }
                            });
                            diagonal_blocks_.resize(rows * rb * rb);
                            source_->blocks(slots, diagonal_blocks_.data());
                        }
                    }
                    bool done = run_rows(rows, budget, cancel, true, [&](size_t f, size_t l, bool) {
                        parallel_for(l - f, 1024, [&](size_t a, size_t b) {
                            for (size_t row = f + a; row < f + b; ++row) {
                                auto first = current_.columns.begin() + current_.starts[row];
                                auto last = current_.columns.begin() + current_.starts[row + 1];
                                auto it = std::lower_bound(first, last, (u32)row);
                                if (it == last || *it != (u32)row)
                                    throw Error(SF_ERR_MG_NOT_POSITIVE_DEFINITE, "a diagonal block is missing", level, (int64_t)row, 0.0);
                                double pivot = 0;
                                const double *block = held ? &diagonal_blocks_[row * rb * rb]
                                                           : current_.block((size_t)(it - current_.columns.begin()));
                                if (!invert_spd(block, rb, &inverse_diagonal_[row * rb * rb], pivot))
                                    throw Error(SF_ERR_MG_NOT_POSITIVE_DEFINITE, "a diagonal block is not positive definite", level,
                                                (int64_t)row, pivot);
                            }
                        });
                    });
                    if (!done) return false;
                    diagonal_blocks_ = std::vector<double>();
                    if (held && !(current_.dofs() <= options_.coarse_dofs || (size_t)level + 1 >= options_.max_levels))
                        // the join's norms come over while the host finds the strong neighbours
                        norms_job_ = std::async(std::launch::async, [this] { return held_block_norms(); });
                    if (accelerator_ && accelerator_->takes(level, current_)) {
                        // the power iteration needs only A and the block inverses: it runs on the
                        // accelerator while the host aggregates
                        on_device_ = true;
                        const double *held = current_.values.empty() && source_ ? source_->device() : nullptr;
                        spectral_job_ = std::async(std::launch::async, [this, level, held] {
                            init_power(false);
                            accelerator_->begin(level, current_, inverse_diagonal_, held);
                            return accelerator_->spectral(power_, (int)POWER_ITERATIONS);
                        });
                    }
                    if (current_.dofs() <= options_.coarse_dofs || (size_t)level + 1 >= options_.max_levels) {
                        is_coarsest_ = true;
                        phase_ = Phase::Spectral;
                    } else {
                        strong_starts_.assign(1, 0);
                        strong_columns_.clear();
                        phase_ = Phase::Strength;
                    }
                    break;
                }
                case Phase::Strength: {
                    // setup_aggregation.rs:12-43, aggregate.rs:53-69
                    if (cursor_ == 0) closest_.clear();
                    bool done = run_rows(rows, budget, cancel, true, [&](size_t f, size_t l, bool) {
                        // each row's closest neighbour and strong count, then the strong columns
                        // written in place at their row's offset
                        const size_t first_row = closest_.size();
                        closest_.resize(first_row + (l - f));
                        strong_starts_.resize(first_row + 1 + (l - f));
                        parallel_for(l - f, 512, [&](size_t a, size_t b) {
                            for (size_t i = a; i < b; ++i) {
                                const size_t row = f + i;
                                double c = std::numeric_limits<double>::max();
                                for (int64_t at = current_.starts[row]; at < current_.starts[row + 1]; ++at) {
                                    size_t col = current_.columns[(size_t)at];
                                    if (col != row) c = std::min(c, dist2(positions_, row, col));
                                }
                                const double limit = reach_ * reach_ * c;
                                int64_t strong = 0;
                                for (int64_t at = current_.starts[row]; at < current_.starts[row + 1]; ++at) {
                                    size_t col = current_.columns[(size_t)at];
                                    if (col != row && dist2(positions_, row, col) <= limit) ++strong;
                                }
                                closest_[first_row + i] = c;
                                strong_starts_[first_row + 1 + i] = strong;
                            }
                        });
                        for (size_t i = 0; i < l - f; ++i) strong_starts_[first_row + 1 + i] += strong_starts_[first_row + i];
                        strong_columns_.resize((size_t)strong_starts_[first_row + (l - f)]);
                        parallel_for(l - f, 512, [&](size_t a, size_t b) {
                            for (size_t i = a; i < b; ++i) {
                                const size_t row = f + i;
                                const double limit = reach_ * reach_ * closest_[first_row + i];
                                size_t to = (size_t)strong_starts_[first_row + i];
                                for (int64_t at = current_.starts[row]; at < current_.starts[row + 1]; ++at) {
                                    size_t col = current_.columns[(size_t)at];
                                    if (col != row && dist2(positions_, row, col) <= limit) strong_columns_[to++] = (u32)col;
                                }
                            }
                        });
                    });
                    if (!done) return false;
                    aggregate_.assign(rows, NONE);
                    seeded_.assign(rows, 0);
                    aggregate_count_ = 0;
                    phase_ = Phase::Seed;
                    break;
                }
                case Phase::Seed: {
                    // aggregate.rs:119-136: a row seeds when it and all its strong neighbours are free
                    if (cursor_ == 0 && accelerator_ && accelerator_->takes(level, current_)) {
                        // on the accelerator the pass is one unit of the rows' budget
                        if (budget == 0) return false;
                        if (cancel()) {
                            cancelled_ = true;
                            throw Error(SF_ERR_MG_CANCELLED, "the multigrid setup was cancelled");
                        }
                        if (accelerator_->seed(level, strong_starts_, strong_columns_, aggregate_, seeded_, aggregate_count_)) {
                            budget -= std::min(budget, rows);
                            phase_ = Phase::Join;
                            break;
                        }
                    }
                    bool done = run_rows(rows, budget, cancel, false, [&](size_t f, size_t l, bool) {
                        for (size_t row = f; row < l; ++row) {
                            const u32 *nb = &strong_columns_[(size_t)strong_starts_[row]];
                            const size_t n = (size_t)(strong_starts_[row + 1] - strong_starts_[row]);
                            if (aggregate_[row] != NONE || n == 0) continue;
                            bool free = true;
                            for (size_t i = 0; i < n && free; ++i) free = aggregate_[nb[i]] == NONE;
                            if (!free) continue;
                            const u32 id = aggregate_count_++;
                            aggregate_[row] = id;
                            seeded_[row] = 1;
                            for (size_t i = 0; i < n; ++i) {
                                aggregate_[nb[i]] = id;
                                seeded_[nb[i]] = 1;
                            }
                        }
                    });
                    if (!done) return false;
                    phase_ = Phase::Join;
                    break;
                }
                case Phase::Join: {
                    // aggregate.rs:141-161: join the strongest seeded neighbour's aggregate. A row
                    // reads only seeded rows' aggregates, which this pass never writes, and writes
                    // only its own, so the rows run in parallel with the serial pass's result.
                    const bool held = source_ && current_.values.empty() && current_.blocks() > 0;
                    if (held && cursor_ == 0) block_norms_ = norms_job_.valid() ? norms_job_.get() : held_block_norms();
                    bool done = run_rows(rows, budget, cancel, true, [&](size_t f, size_t l, bool) {
                        parallel_for(l - f, 1024, [&](size_t a, size_t b) {
                            for (size_t row = f + a; row < f + b; ++row) {
                                if (aggregate_[row] != NONE) continue;
                                const u32 *nb = &strong_columns_[(size_t)strong_starts_[row]];
                                const u32 *ne = nb + (strong_starts_[row + 1] - strong_starts_[row]);
                                double best = 0.0;
                                u32 chosen = NONE;
                                for (int64_t at = current_.starts[row]; at < current_.starts[row + 1]; ++at) {
                                    const u32 col = current_.columns[(size_t)at];
                                    if (!std::binary_search(nb, ne, col) || !seeded_[col]) continue;
                                    const double norm = held ? block_norms_[(size_t)at] : frobenius(current_.block((size_t)at), rb * current_.cb);
                                    if (norm > best) {
                                        best = norm;
                                        chosen = aggregate_[col];
                                    }
                                }
                                if (chosen != NONE) aggregate_[row] = chosen;
                            }
                        });
                    });
                    if (!done) return false;
                    block_norms_ = std::vector<double>();
                    phase_ = Phase::Sweep;
                    break;
                }
                case Phase::Sweep: {
                    // aggregate.rs:165-177: every remaining row starts an aggregate
                    bool done = run_rows(rows, budget, cancel, false, [&](size_t f, size_t l, bool) {
                        for (size_t row = f; row < l; ++row) {
                            if (aggregate_[row] != NONE) continue;
                            const u32 id = aggregate_count_++;
                            aggregate_[row] = id;
                            for (int64_t at = strong_starts_[row]; at < strong_starts_[row + 1]; ++at) {
                                const u32 n = strong_columns_[(size_t)at];
                                if (aggregate_[n] == NONE) aggregate_[n] = id;
                            }
                        }
                    });
                    if (!done) return false;
                    phase_ = Phase::Group;
                    break;
                }
                case Phase::Group: {
                    // setup_aggregation.rs:165-228
                    const size_t count = aggregate_count_;
                    const bool coarsens = (double)(count * MODES) <= MAX_COARSE_FRACTION * (double)current_.dofs();
                    if (!coarsens && retries_ < MAX_RETRIES) {
                        retries_ += 1;
                        reach_ *= RETRY_WIDENING;
                        strong_starts_.assign(1, 0);
                        strong_columns_.clear();
                        phase_ = Phase::Strength;
                        break;
                    }
                    if (!coarsens || count == 0) {
                        is_coarsest_ = true;
                        phase_ = Phase::Spectral;
                        break;
                    }
                    member_starts_.assign(count + 1, 0);
                    for (u32 a : aggregate_) member_starts_[(size_t)a + 1] += 1;
                    for (size_t i = 0; i < count; ++i) member_starts_[i + 1] += member_starts_[i];
                    std::vector<int64_t> fill(member_starts_.begin(), member_starts_.end() - 1);
                    members_.assign(rows, 0);
                    for (size_t row = 0; row < rows; ++row) members_[(size_t)fill[aggregate_[row]]++] = (u32)row;
                    coarse_positions_.assign(count, Vec3{{0, 0, 0}});
                    for (size_t a = 0; a < count; ++a) {
                        for (int64_t m = member_starts_[a]; m < member_starts_[a + 1]; ++m)
                            for (int x = 0; x < 3; ++x) coarse_positions_[a][x] += positions_[members_[(size_t)m]][x];
                        const double n = (double)(member_starts_[a + 1] - member_starts_[a]);
                        for (int x = 0; x < 3; ++x) coarse_positions_[a][x] /= n;
                    }
                    tentative_values_.assign(rows * rb * MODES, 0.0);
                    coarse_null_.assign(count * 36, 0.0);
                    phase_ = Phase::Tentative;
                    break;
                }
                case Phase::Tentative: {
                    // setup.rs:327-386, transfer.rs:19-68: two-pass Gram-Schmidt per aggregate
                    bool done = run_rows(aggregate_count_, budget, cancel, true, [&](size_t f, size_t l, bool) {
                        parallel_for(l - f, 256, [&](size_t a0, size_t a1) {
                            std::vector<double> local;
                            for (size_t agg = f + a0; agg < f + a1; ++agg) {
                                const u32 *mem = &members_[(size_t)member_starts_[agg]];
                                const size_t cnt = (size_t)(member_starts_[agg + 1] - member_starts_[agg]);
                                const size_t n = cnt * rb;
                                local.assign(n * MODES, 0.0);
                                for (size_t m = 0; m < cnt; ++m)
                                    std::copy(&near_null_[(size_t)mem[m] * rb * MODES], &near_null_[(size_t)mem[m] * rb * MODES] + rb * MODES,
                                              &local[m * rb * MODES]);
                                double *coarse = &coarse_null_[agg * 36];
                                auto colnorm = [&](size_t c) {
                                    double s = 0.0;
                                    for (size_t r = 0; r < n; ++r) s += local[r * MODES + c] * local[r * MODES + c];
                                    return std::sqrt(s);
                                };
                                for (size_t c = 0; c < MODES; ++c) {
                                    const double original = colnorm(c);
                                    for (int pass = 0; pass < 2; ++pass)
                                        for (size_t pc = 0; pc < c; ++pc) {
                                            double proj = 0.0;
                                            for (size_t r = 0; r < n; ++r) proj += local[r * MODES + pc] * local[r * MODES + c];
                                            coarse[pc * MODES + c] += proj;
                                            for (size_t r = 0; r < n; ++r) local[r * MODES + c] -= proj * local[r * MODES + pc];
                                        }
                                    const double rem = colnorm(c);
                                    if (original == 0.0 || rem <= RANK_TOLERANCE * original) {
                                        for (size_t r = 0; r < n; ++r) local[r * MODES + c] = 0.0;
                                        coarse[c * MODES + c] = 0.0;
                                    } else {
                                        for (size_t r = 0; r < n; ++r) local[r * MODES + c] /= rem;
                                        coarse[c * MODES + c] = rem;
                                    }
                                }
                                for (size_t m = 0; m < cnt; ++m)
                                    std::copy(&local[m * rb * MODES], &local[m * rb * MODES] + rb * MODES,
                                              &tentative_values_[(size_t)mem[m] * rb * MODES]);
                            }
                        });
                    });
                    if (!done) return false;
                    phase_ = Phase::Spectral;
                    break;
                }
                case Phase::Spectral: {
                    // setup.rs:390-449
                    const size_t dofs = current_.dofs();
                    if (power_done_ == 0 && !spectral_job_.valid() && power_.size() != dofs) init_power(true);
                    if (power_done_ == 0 && spectral_job_.valid()) {
                        // the iteration started beside the aggregation, charged as the host's
                        if (budget == 0) return false;
                        if (cancel()) {
                            cancelled_ = true;
                            throw Error(SF_ERR_MG_CANCELLED, "the multigrid setup was cancelled");
                        }
                        lambda_max_ = spectral_job_.get();
                        power_done_ = POWER_ITERATIONS;
                        budget = budget > rows * POWER_ITERATIONS ? budget - rows * POWER_ITERATIONS : 0;
                    }
                    if (power_done_ < POWER_ITERATIONS) ensure_values();
                    while (power_done_ < POWER_ITERATIONS) {
                        if (budget == 0) return false;
                        if (cancel()) {
                            cancelled_ = true;
                            throw Error(SF_ERR_MG_CANCELLED, "the multigrid setup was cancelled");
                        }
                        spectral_iteration();
                        power_done_ += 1;
                        budget = budget > rows ? budget - rows : 0;
                    }
                    power_done_ = 0;
                    power_.clear();
                    scratch_.clear();
                    scratch2_.clear();
                    lambda_max_ *= SPECTRAL_MARGIN;
                    if (is_coarsest_) {
                        if (current_.dofs() <= DENSE_LIMIT) {
                            // setup.rs:451-468
                            ensure_values();
                            HostLevel l;
                            const size_t n = current_.dofs();
                            l.factor.assign(n * n, 0.0);
                            for (size_t row = 0; row < rows; ++row)
                                for (int64_t at = current_.starts[row]; at < current_.starts[row + 1]; ++at) {
                                    const size_t col = current_.columns[(size_t)at];
                                    const double *b = current_.block((size_t)at);
                                    for (size_t p = 0; p < rb; ++p)
                                        for (size_t q = 0; q < current_.cb; ++q) l.factor[(row * rb + p) * n + col * current_.cb + q] = b[p * current_.cb + q];
                                }
                            levels_.push_back(std::move(l));
                            release_accelerator_level();
                            cursor_ = 0;
                            phase_ = Phase::Factor;
                        } else {
                            push_final_level();
                            phase_ = Phase::Done;
                        }
                        break;
                    }
                    // the tentative prolongator: one block per row in its aggregate's column
                    prolongator_ = HostBsr<double>();
                    prolongator_.rows = rows;
                    prolongator_.cols = aggregate_count_;
                    prolongator_.rb = rb;
                    prolongator_.cb = MODES;
                    phase_ = Phase::Filter;
                    break;
                }
                case Phase::Filter: {
                    // setup_aggregation.rs:48-84, aggregate.rs:80-96 (coarse levels only)
                    if (level == 0) {
                        phase_ = Phase::Smooth;
                        break;
                    }
                    if (cursor_ == 0) {
                        filter_starts_.assign(1, 0);
                        filter_columns_.clear();
                    }
                    bool done = run_rows(rows, budget, cancel, true, [&](size_t f, size_t l, bool) {
                        std::vector<RowOut> out(l - f);
                        parallel_for(l - f, 512, [&](size_t a, size_t b) {
                            for (size_t i = a; i < b; ++i) {
                                const size_t row = f + i;
                                double scale = closest_[row];
                                for (int64_t at = current_.starts[row]; at < current_.starts[row + 1]; ++at) {
                                    size_t col = current_.columns[(size_t)at];
                                    if (col != row) scale = std::max(scale, closest_[col]);
                                }
                                const double limit = reach_ * reach_ * scale;
                                for (int64_t at = current_.starts[row]; at < current_.starts[row + 1]; ++at) {
                                    size_t col = current_.columns[(size_t)at];
                                    if (col != row && dist2(positions_, row, col) <= limit) out[i].cols.push_back((u32)col);
                                }
                            }
                        });
                        for (auto &o : out) {
                            filter_columns_.insert(filter_columns_.end(), o.cols.begin(), o.cols.end());
                            filter_starts_.push_back((int64_t)filter_columns_.size());
                        }
                    });
                    if (!done) return false;
                    phase_ = Phase::Smooth;
                    break;
                }
                case Phase::Smooth: {
                    // setup.rs:527-573, transfer.rs:108-152: P = T - omega D^-1 A_f T
                    const double omega = 4.0 / (3.0 * lambda_max_);
                    const bool use_filter = !filter_starts_.empty();
                    const std::vector<int64_t> &ns = use_filter ? filter_starts_ : strong_starts_;
                    const std::vector<u32> &nc = use_filter ? filter_columns_ : strong_columns_;
                    const size_t tb = rb * MODES;
                    bool on_accelerator = false;
                    if (on_device_ && cursor_ == 0) {
                        if (budget == 0) return false;
                        if (cancel()) {
                            cancelled_ = true;
                            throw Error(SF_ERR_MG_CANCELLED, "the multigrid setup was cancelled");
                        }
                        on_accelerator = device_p_ = accelerator_->smooth(aggregate_, aggregate_count_, ns, nc, tentative_values_, omega, prolongator_);
                        if (on_accelerator) {
                            p_source_ = accelerator_->prolongator_values();
                            budget = budget > rows ? budget - rows : 0;
                        } else if (StageClock::on()) {
                            std::fprintf(stderr, "[setup] level %d: the smoothing falls back to the host\n", level);
                        }
                    }
                    if (!on_accelerator) ensure_values();
                    bool done = on_accelerator || run_rows(rows, budget, cancel, true, [&](size_t f, size_t l, bool) {
                        std::vector<RowOut> out(l - f);
                        parallel_for(l - f, 256, [&](size_t a, size_t b) {
                            RowAccumulator acc(aggregate_count_, tb);
                            std::vector<double> scaled(tb);
                            for (size_t i = a; i < b; ++i) {
                                const size_t row = f + i;
                                acc.clear();
                                const u32 t_col = aggregate_[row];
                                acc.at(t_col);
                                const u32 *nb = &nc[(size_t)ns[row]];
                                const u32 *ne = nb + (ns[row + 1] - ns[row]);
                                for (int64_t at = current_.starts[row]; at < current_.starts[row + 1]; ++at) {
                                    const u32 col = current_.columns[(size_t)at];
                                    const size_t src = ((size_t)col == row || std::binary_search(nb, ne, col)) ? col : row;
                                    gemm_acc(acc.at(aggregate_[src]), current_.block((size_t)at), rb, current_.cb,
                                             &tentative_values_[src * tb], MODES);
                                }
                                acc.sort();
                                const double *dinv = &inverse_diagonal_[row * rb * rb];
                                const double *t_block = &tentative_values_[row * tb];
                                for (u32 col : acc.touched) {
                                    std::fill(scaled.begin(), scaled.end(), 0.0);
                                    gemm_acc(scaled.data(), dinv, rb, rb, acc.get(col), MODES);
                                    out[i].cols.push_back(col);
                                    for (size_t k = 0; k < tb; ++k) out[i].vals.push_back((col == t_col ? t_block[k] : 0.0) - omega * scaled[k]);
                                }
                            }
                        });
                        append_rows(prolongator_, out);
                    });
                    if (!done) return false;
                    tentative_values_.clear();
                    filter_starts_.clear();
                    filter_columns_.clear();
                    a_times_p_ = HostBsr<double>();
                    a_times_p_.rows = rows;
                    a_times_p_.cols = aggregate_count_;
                    a_times_p_.rb = rb;
                    a_times_p_.cb = MODES;
                    phase_ = Phase::Product;
                    break;
                }
                case Phase::Product: {
                    // setup.rs:575-612, transfer.rs:81-95 and 155-169: A P row by row
                    const size_t tb = rb * MODES;
                    bool on_accelerator = false;
                    if (device_p_ && cursor_ == 0) {
                        if (budget == 0) return false;
                        if (cancel()) {
                            cancelled_ = true;
                            throw Error(SF_ERR_MG_CANCELLED, "the multigrid setup was cancelled");
                        }
                        on_accelerator = device_ap_ = accelerator_->product(aggregate_count_);
                        if (on_accelerator) budget = budget > rows ? budget - rows : 0;
                        else if (StageClock::on()) std::fprintf(stderr, "[setup] level %d: A P falls back to the host\n", level);
                    }
                    if (!on_accelerator) {
                        ensure_values();
                        ensure_prolongator_values();
                    }
                    bool done = on_accelerator || run_rows(rows, budget, cancel, true, [&](size_t f, size_t l, bool) {
                        std::vector<RowOut> out(l - f);
                        parallel_for(l - f, 256, [&](size_t a, size_t b) {
                            RowAccumulator acc(aggregate_count_, tb);
                            for (size_t i = a; i < b; ++i) {
                                const size_t row = f + i;
                                acc.clear();
                                for (int64_t at = current_.starts[row]; at < current_.starts[row + 1]; ++at) {
                                    const size_t col = current_.columns[(size_t)at];
                                    for (int64_t in = prolongator_.starts[col]; in < prolongator_.starts[col + 1]; ++in)
                                        gemm_acc(acc.at(prolongator_.columns[(size_t)in]), current_.block((size_t)at), rb, current_.cb,
                                                 prolongator_.block((size_t)in), MODES);
                                }
                                acc.sort();
                                for (u32 c : acc.touched) {
                                    out[i].cols.push_back(c);
                                    out[i].vals.insert(out[i].vals.end(), acc.get(c), acc.get(c) + tb);
                                }
                            }
                        });
                        append_rows(a_times_p_, out);
                    });
                    if (!done) return false;
                    // transfer.rs:173-194: the prolongator by columns, fine rows ascending
                    const size_t count = aggregate_count_;
                    pcol_starts_.assign(count + 1, 0);
                    for (u32 c : prolongator_.columns) pcol_starts_[(size_t)c + 1] += 1;
                    for (size_t i = 0; i < count; ++i) pcol_starts_[i + 1] += pcol_starts_[i];
                    std::vector<int64_t> fill(pcol_starts_.begin(), pcol_starts_.end() - 1);
                    pcol_entries_.assign(prolongator_.columns.size(), {0, 0});
                    for (size_t row = 0; row < rows; ++row)
                        for (int64_t at = prolongator_.starts[row]; at < prolongator_.starts[row + 1]; ++at)
                            pcol_entries_[(size_t)fill[prolongator_.columns[(size_t)at]]++] = {(u32)row, at};
                    coarse_ = HostBsr<double>();
                    coarse_.rows = coarse_.cols = count;
                    coarse_.rb = coarse_.cb = MODES;
                    phase_ = Phase::Galerkin;
                    break;
                }
                case Phase::Galerkin: {
                    // setup.rs:614-647, transfer.rs:204-238: P^T (A P) by coarse rows
                    const size_t count = aggregate_count_;
                    bool on_accelerator = false;
                    if (device_ap_ && cursor_ == 0) {
                        if (budget == 0) return false;
                        if (cancel()) {
                            cancelled_ = true;
                            throw Error(SF_ERR_MG_CANCELLED, "the multigrid setup was cancelled");
                        }
                        on_accelerator = accelerator_->galerkin((u32)count, pcol_starts_, pcol_entries_, coarse_);
                        if (on_accelerator) {
                            c_source_ = accelerator_->coarse_values();
                            budget = budget > count ? budget - count : 0;
                        }
                        else if (StageClock::on()) std::fprintf(stderr, "[setup] level %d: the Galerkin product falls back to the host\n", level);
                    }
                    if (!on_accelerator) {
                        ensure_prolongator_values();
                        if (device_ap_ && a_times_p_.columns.empty() && a_times_p_.starts.size() != a_times_p_.rows + 1)
                            accelerator_->product_values(a_times_p_);
                    }
                    bool done = on_accelerator || run_rows(count, budget, cancel, true, [&](size_t f, size_t l, bool) {
                        std::vector<RowOut> out(l - f);
                        parallel_for(l - f, 64, [&](size_t a, size_t b) {
                            RowAccumulator acc(count, MODES * MODES);
                            std::vector<double> tr(rb * MODES);
                            for (size_t i = a; i < b; ++i) {
                                const size_t c = f + i;
                                acc.clear();
                                for (int64_t e = pcol_starts_[c]; e < pcol_starts_[c + 1]; ++e) {
                                    const size_t fine_row = pcol_entries_[(size_t)e].first;
                                    const double *blk = prolongator_.block((size_t)pcol_entries_[(size_t)e].second);
                                    for (size_t p = 0; p < rb; ++p)
                                        for (size_t q = 0; q < MODES; ++q) tr[q * rb + p] = blk[p * MODES + q];
                                    for (int64_t in = a_times_p_.starts[fine_row]; in < a_times_p_.starts[fine_row + 1]; ++in)
                                        gemm_acc(acc.at(a_times_p_.columns[(size_t)in]), tr.data(), MODES, rb, a_times_p_.block((size_t)in), MODES);
                                }
                                acc.sort();
                                for (u32 col : acc.touched) {
                                    out[i].cols.push_back(col);
                                    out[i].vals.insert(out[i].vals.end(), acc.get(col), acc.get(col) + MODES * MODES);
                                }
                            }
                        });
                        append_rows(coarse_, out);
                    });
                    if (!done) return false;
                    finish_level();
                    break;
                }
                case Phase::Factor: {
                    // setup.rs:470-504, block.rs:327-348: one column of the dense Cholesky per step
                    HostLevel &l = levels_.back();
                    const size_t n = current_.dofs();
                    double *a = l.factor.data();
                    if (cursor_ == 0 && accelerator_) {
                        // on the accelerator the factor is one unit of the budget
                        if (budget == 0) return false;
                        if (cancel()) {
                            cancelled_ = true;
                            throw Error(SF_ERR_MG_CANCELLED, "the multigrid setup was cancelled");
                        }
                        if (accelerator_->factor(level, n, a)) {
                            budget -= std::min(budget, n);
                            cursor_ = n;
                        }
                    }
                    while (cursor_ < n) {
                        if (budget == 0) return false;
                        if (cancel()) {
                            cancelled_ = true;
                            throw Error(SF_ERR_MG_CANCELLED, "the multigrid setup was cancelled");
                        }
                        const size_t j = cursor_;
                        double s = a[j * n + j];
                        for (size_t k = 0; k < j; ++k) s -= a[j * n + k] * a[j * n + k];
                        if (!std::isfinite(s) || s <= 0.0)
                            throw Error(SF_ERR_MG_NOT_POSITIVE_DEFINITE, "the coarse matrix is not positive definite", level, (int64_t)j, s);
                        a[j * n + j] = std::sqrt(s);
                        parallel_for(n - j - 1, 64, [&](size_t f, size_t e) {
                            for (size_t i = j + 1 + f; i < j + 1 + e; ++i) {
                                double t = a[i * n + j];
                                for (size_t k = 0; k < j; ++k) t -= a[i * n + k] * a[j * n + k];
                                a[i * n + j] = t / a[j * n + j];
                            }
                        });
                        cursor_ += 1;
                        budget = budget > n ? budget - n : 0;
                    }
                    cursor_ = 0;
                    l.has_factor = true;
                    const auto range = std::make_pair(lambda_max_ / options_.range, lambda_max_);
                    l.lambda_min = range.first;
                    l.lambda_max = range.second;
                    l.a = std::move(current_);
                    l.inverse_diagonal = std::move(inverse_diagonal_);
                    phase_ = Phase::Done;
                    break;
                }
            }
        }
    } catch (Error &e) {
        if (e.status != SF_ERR_MG_CANCELLED) {
            failed_ = true;
            failure_ = e;
        }
        throw;
    }
}

void MultigridSetup::release_accelerator_level() {
    if (on_device_) accelerator_->end_level();
    on_device_ = device_p_ = device_ap_ = false;
}

std::vector<double> MultigridSetup::held_block_norms() const {
    // the host's Frobenius norms from the sums of squares summed on the device in its order
    std::vector<double> norms(current_.blocks());
    source_->block_squares(norms.data());
    parallel_for(norms.size(), 65536, [&](size_t a, size_t b) {
        for (size_t i = a; i < b; ++i) norms[i] = std::sqrt(norms[i]);
    });
    return norms;
}

void MultigridSetup::init_power(bool host_iteration) {
    const size_t dofs = current_.dofs();
    power_.resize(dofs);
    for (size_t i = 0; i < dofs; ++i) {
        const u64 mixed = ((u64)i + 1) * 0x9E3779B97F4A7C15ull;  // block.rs:379-382
        power_[i] = 0.5 + (double)(mixed >> 11) / (double)(1ull << 53);
    }
    double s = 0.0;
    for (double v : power_) s += v * v;
    const double norm = std::sqrt(s);
    for (double &v : power_) v /= norm;
    if (host_iteration) {
        scratch_.assign(dofs, 0.0);
        scratch2_.assign(dofs, 0.0);
    }
}

void MultigridSetup::ensure_values() {
    if (!source_ || !current_.values.empty() || current_.blocks() == 0) return;
    current_.values.resize(current_.blocks() * current_.block_len());
    source_->values(current_.values.data());
}

void MultigridSetup::ensure_prolongator_values() {
    if (!p_source_ || !prolongator_.values.empty() || prolongator_.blocks() == 0) return;
    prolongator_.values.resize(prolongator_.blocks() * prolongator_.block_len());
    p_source_->values(prolongator_.values.data());
}

void HostLevel::fetch_values() const {
    if (source && a.values.empty() && a.blocks() > 0) {
        a.values.resize(a.blocks() * a.block_len());
        source->values(a.values.data());
    }
    if (prolongator_source && prolongator.values.empty() && prolongator.blocks() > 0) {
        prolongator.values.resize(prolongator.blocks() * prolongator.block_len());
        prolongator_source->values(prolongator.values.data());
    }
}

void MultigridSetup::push_final_level() {
    release_accelerator_level();
    HostLevel l;
    if (current_.values.empty()) l.source = source_;
    source_.reset();
    l.lambda_min = lambda_max_ / options_.range;
    l.lambda_max = lambda_max_;
    l.a = std::move(current_);
    l.inverse_diagonal = std::move(inverse_diagonal_);
    levels_.push_back(std::move(l));
}

void MultigridSetup::finish_level() {
    std::shared_ptr<const ValueSource> p_source = prolongator_.values.empty() ? p_source_ : nullptr;
    p_source_.reset();
    release_accelerator_level();
    // setup.rs:650-695: a unit diagonal on coarse values no aggregate direction reaches (the
    // device's Galerkin product applies it as it writes)
    const size_t dim = coarse_.rb;
    for (size_t row = 0; row < (coarse_.values.empty() ? 0 : coarse_.rows); ++row)
        for (int64_t at = coarse_.starts[row]; at < coarse_.starts[row + 1]; ++at) {
            if (coarse_.columns[(size_t)at] != row) continue;
            for (size_t d = 0; d < dim; ++d) {
                double &v = coarse_.values[(size_t)at * dim * dim + d * dim + d];
                if (v == 0.0) v = 1.0;
            }
        }
    HostLevel l;
    if (current_.values.empty()) l.source = source_;
    source_.reset();
    l.lambda_min = lambda_max_ / options_.range;
    l.lambda_max = lambda_max_;
    l.a = std::move(current_);
    l.inverse_diagonal = std::move(inverse_diagonal_);
    l.has_prolongator = true;
    l.prolongator = std::move(prolongator_);
    l.prolongator_source = std::move(p_source);
    l.aggregates = std::move(aggregate_);
    levels_.push_back(std::move(l));
    current_ = std::move(coarse_);
    coarse_ = HostBsr<double>();
    source_ = current_.values.empty() ? std::move(c_source_) : nullptr;
    c_source_.reset();
    a_times_p_ = HostBsr<double>();
    pcol_starts_.clear();
    pcol_entries_.clear();
    near_null_ = std::move(coarse_null_);
    positions_ = std::move(coarse_positions_);
    reach_ = options_.reach;
    retries_ = 0;
    is_coarsest_ = false;
    cursor_ = 0;
    phase_ = Phase::Diagonal;
}

}  // namespace sf
