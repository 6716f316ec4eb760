// refcpu multigrid: TEST INFRASTRUCTURE (the oracle). Restates
// crates/solidfem-solve/src/multigrid/{block,aggregate,transfer,setup,setup_aggregation,cycle}.rs
// in the reference's floating-point order. Setup is run unsliced: the reference promises that
// slicing never changes the bits (crates/solidfem-solve/src/lib.rs:15-19).
#pragma once

#include "refcpu_core.hpp"
#include "refcpu_parallel.hpp"

namespace refcpu {

// multigrid/mod.rs:124
constexpr size_t MODES = 6;
// setup.rs:28-45
constexpr size_t DENSE_LIMIT = 3000;
constexpr double MAX_COARSE_FRACTION = 0.7;
constexpr double RETRY_WIDENING = 1.5;
constexpr size_t MAX_RETRIES = 3;
constexpr size_t POWER_ITERATIONS = 20;
constexpr double SPECTRAL_MARGIN = 1.15;
// transfer.rs:12
constexpr double RANK_TOLERANCE = 1e-10;
constexpr u32 NO_AGGREGATE = 0xffffffffu;

enum class Precision { Double = 0, SingleStorage = 1, Single = 2 };

struct MultigridOptions {
    double neighbour_reach = 3.0;
    size_t coarse_dofs = 1500;
    size_t max_levels = 12;
    size_t smoother_degree = 3;
    double smoother_range = 30.0;
    Precision precision = Precision::SingleStorage;
    size_t threads = 1;  // MultigridOptions::threads (mod.rs:162-167): threads of the cycle's row products
};

struct SetupError : std::runtime_error {
    int level;
    size_t block;
    double pivot;
    SetupError(int l, size_t b, double p)
        : std::runtime_error("multigrid block not positive definite"), level(l), block(b), pivot(p) {}
};

// The value-type conversion of multigrid/scalar.rs:12-65: widen = from_f64(to_f64()).
template <typename V, typename S>
inline V widen(S s) {
    return (V)(double)s;
}

// block.rs:13-180
template <typename T>
struct BlockCsr {
    size_t rows = 0, row_block = 0, col_block = 0;
    std::vector<size_t> starts{0};
    std::vector<u32> columns;
    std::vector<T> values;

    size_t block_len() const { return row_block * col_block; }
    size_t dofs() const { return rows * row_block; }
    const T *block(size_t at) const { return &values[at * block_len()]; }

    static BlockCsr with_shape(size_t rows, size_t rb, size_t cb) {
        BlockCsr m;
        m.rows = rows;
        m.row_block = rb;
        m.col_block = cb;
        return m;
    }

    template <typename U>
    BlockCsr<U> into_precision() const {
        BlockCsr<U> m;
        m.rows = rows;
        m.row_block = row_block;
        m.col_block = col_block;
        m.starts = starts;
        m.columns = columns;
        m.values.resize(values.size());
        for (size_t i = 0; i < values.size(); ++i) m.values[i] = (U)values[i];
        return m;
    }

    // block.rs:136-154
    template <typename V>
    void multiply_rows(size_t first, size_t last, const V *x, V *y) const {
        size_t rb = row_block, cb = col_block;
        for (size_t row = first; row < last; ++row) {
            V *out = y + (row - first) * rb;
            for (size_t p = 0; p < rb; ++p) out[p] = (V)0;
            for (size_t at = starts[row]; at < starts[row + 1]; ++at) {
                const T *b = block(at);
                size_t col = (size_t)columns[at] * cb;
                for (size_t p = 0; p < rb; ++p) {
                    V sum = out[p];
                    for (size_t q = 0; q < cb; ++q) sum += widen<V>(b[p * cb + q]) * x[col + q];
                    out[p] = sum;
                }
            }
        }
    }
    template <typename V>
    void multiply(const V *x, V *y) const {
        multiply_rows(0, rows, x, y);
    }

    // block.rs:157-172
    template <typename V>
    void multiply_transpose(const V *x, V *y, size_t y_len) const {
        size_t rb = row_block, cb = col_block;
        for (size_t i = 0; i < y_len; ++i) y[i] = (V)0;
        for (size_t row = 0; row < rows; ++row)
            for (size_t at = starts[row]; at < starts[row + 1]; ++at) {
                const T *b = block(at);
                size_t col = (size_t)columns[at] * cb;
                for (size_t p = 0; p < rb; ++p) {
                    V value = x[row * rb + p];
                    for (size_t q = 0; q < cb; ++q) y[col + q] += widen<V>(b[p * cb + q]) * value;
                }
            }
    }

    size_t memory_bytes() const {
        return starts.size() * sizeof(size_t) + columns.size() * sizeof(u32) + values.size() * sizeof(T);
    }
};

// block.rs:51-99
inline BlockCsr<double> block_from_nodal(const CsrMatrix &m) {
    if (m.size % 3 != 0) throw std::runtime_error("matrix does not group into 3x3 blocks");
    size_t rows = m.size / 3;
    auto b = BlockCsr<double>::with_shape(rows, 3, 3);
    b.columns.reserve(m.columns.size() / 9);
    b.values.reserve(m.columns.size());
    for (size_t node = 0; node < rows; ++node) {
        size_t s[3], e[3];
        for (int r = 0; r < 3; ++r) {
            s[r] = m.row_starts[3 * node + r];
            e[r] = m.row_starts[3 * node + r + 1];
        }
        size_t len = e[0] - s[0];
        for (int r = 1; r < 3; ++r)
            if (e[r] - s[r] != len || !std::equal(&m.columns[s[0]], &m.columns[s[0]] + len, &m.columns[s[r]]))
                throw std::runtime_error("matrix rows do not share a block pattern");
        if (len % 3 != 0) throw std::runtime_error("matrix row is not whole blocks");
        for (size_t at = 0; at < len; at += 3) {
            u32 c = m.columns[s[0] + at];
            if (c % 3 != 0 || m.columns[s[0] + at + 1] != c + 1 || m.columns[s[0] + at + 2] != c + 2)
                throw std::runtime_error("matrix row is not whole blocks");
            b.columns.push_back(c / 3);
        }
        for (size_t at = 0; at < len / 3; ++at)
            for (int r = 0; r < 3; ++r)
                for (int q = 0; q < 3; ++q) b.values.push_back(m.values[s[r] + 3 * at + q]);
        b.starts.push_back(b.columns.size());
    }
    return b;
}

// block.rs:186-224; returns false and the pivot on failure
inline bool invert_spd_block(const double *input, size_t n, double *output, double &pivot) {
    double lower[36] = {0};
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j <= i; ++j) {
            double sum = input[i * n + j];
            for (size_t k = 0; k < j; ++k) sum -= lower[i * n + k] * lower[j * n + k];
            if (i == j) {
                if (!std::isfinite(sum) || sum <= 0.0) {
                    pivot = sum;
                    return false;
                }
                lower[i * n + j] = std::sqrt(sum);
            } else {
                lower[i * n + j] = sum / lower[j * n + j];
            }
        }
    double inverse[36] = {0};
    for (size_t column = 0; column < n; ++column)
        for (size_t i = column; i < n; ++i) {
            double sum = (i == column) ? 1.0 : 0.0;
            for (size_t k = column; k < i; ++k) sum -= lower[i * n + k] * inverse[k * n + column];
            inverse[i * n + column] = sum / lower[i * n + i];
        }
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) {
            double sum = 0.0;
            for (size_t k = std::max(i, j); k < n; ++k) sum += inverse[k * n + i] * inverse[k * n + j];
            output[i * n + j] = sum;
        }
    return true;
}

// block.rs:229-245
inline void add_product(double *out, const double *a, size_t a_rows, size_t middle, const double *b, size_t b_cols) {
    for (size_t row = 0; row < a_rows; ++row)
        for (size_t mid = 0; mid < middle; ++mid) {
            double scale = a[row * middle + mid];
            for (size_t col = 0; col < b_cols; ++col) out[row * b_cols + col] += scale * b[mid * b_cols + col];
        }
}

// block.rs:249-302
struct Accumulator {
    std::vector<u32> slot;
    std::vector<u32> touched;
    std::vector<double> data;
    size_t block_len;
    Accumulator(size_t columns, size_t bl) : slot(columns, 0xffffffffu), block_len(bl) {}
    void begin() {
        for (u32 c : touched) slot[c] = 0xffffffffu;
        touched.clear();
        data.clear();
    }
    double *entry(u32 column) {
        u32 at = slot[column];
        if (at == 0xffffffffu) {
            at = (u32)touched.size();
            slot[column] = at;
            touched.push_back(column);
            data.resize(data.size() + block_len, 0.0);
        }
        return &data[(size_t)at * block_len];
    }
    template <typename F>
    void drain_sorted(F emit) {
        std::sort(touched.begin(), touched.end());
        for (u32 c : touched) emit(c, &data[(size_t)slot[c] * block_len]);
    }
};

// block.rs:307-374
struct DenseCholesky {
    size_t n = 0;
    std::vector<double> factor;
    size_t next = 0;
    // returns 1 done, 0 more, -1 failure (pivot and column set)
    int factor_column(size_t &col, double &pivot) {
        size_t j = next;
        if (j == n) return 1;
        for (size_t i = j; i < n; ++i) {
            double sum = factor[i * n + j];
            for (size_t k = 0; k < j; ++k) sum -= factor[i * n + k] * factor[j * n + k];
            if (i == j) {
                if (!std::isfinite(sum) || sum <= 0.0) {
                    col = j;
                    pivot = sum;
                    return -1;
                }
                factor[j * n + j] = std::sqrt(sum);
            } else {
                factor[i * n + j] = sum / factor[j * n + j];
            }
        }
        next += 1;
        return next == n ? 1 : 0;
    }
    void solve(double *x) const {
        for (size_t i = 0; i < n; ++i) {
            double sum = x[i];
            for (size_t k = 0; k < i; ++k) sum -= factor[i * n + k] * x[k];
            x[i] = sum / factor[i * n + i];
        }
        for (size_t i = n; i-- > 0;) {
            double sum = x[i];
            for (size_t k = i + 1; k < n; ++k) sum -= factor[k * n + i] * x[k];
            x[i] = sum / factor[i * n + i];
        }
    }
    size_t memory_bytes() const { return factor.size() * sizeof(double); }
};

// block.rs:379-382
inline double start_value(size_t index) {
    u64 mixed = ((u64)index + 1) * 0x9E3779B97F4A7C15ull;
    return 0.5 + (double)(mixed >> 11) / (double)(1ull << 53);
}

// block.rs:384-396
inline void apply_block_diagonal(const double *inverse, size_t block, const double *x, double *y, size_t len) {
    size_t rows = len / block;
    for (size_t row = 0; row < rows; ++row) {
        const double *d = inverse + row * block * block;
        for (size_t p = 0; p < block; ++p) {
            double sum = 0.0;
            for (size_t q = 0; q < block; ++q) sum += d[p * block + q] * x[row * block + q];
            y[row * block + p] = sum;
        }
    }
}

// block.rs:398-411
inline double norm2(const double *x, size_t n) {
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) sum += x[i] * x[i];
    return std::sqrt(sum);
}

// ---------------------------------------------------------------------------------------------
// Aggregation: aggregate.rs
// ---------------------------------------------------------------------------------------------

inline double block_norm(const double *b, size_t len) {
    double sum = 0.0;
    for (size_t i = 0; i < len; ++i) sum += b[i] * b[i];
    return std::sqrt(sum);
}

inline double squared_distance(const std::vector<Vec3> &pos, size_t row, size_t col) {
    const Vec3 &h = pos[row], &t = pos[col];
    double dx = t[0] - h[0], dy = t[1] - h[1], dz = t[2] - h[2];
    return dx * dx + dy * dy + dz * dz;
}

// aggregate.rs:34-48
inline void push_within(const BlockCsr<double> &m, const std::vector<Vec3> &pos, size_t row, double limit,
                        std::vector<u32> &cols) {
    for (size_t at = m.starts[row]; at < m.starts[row + 1]; ++at) {
        size_t c = m.columns[at];
        if (c != row && squared_distance(pos, row, c) <= limit) cols.push_back((u32)c);
    }
}

// aggregate.rs:53-69
inline double strong_row(const BlockCsr<double> &m, const std::vector<Vec3> &pos, double reach, size_t row,
                         std::vector<u32> &cols) {
    double closest = 1.7976931348623157e308;
    for (size_t at = m.starts[row]; at < m.starts[row + 1]; ++at) {
        size_t c = m.columns[at];
        if (c != row) closest = std::min(closest, squared_distance(pos, row, c));
    }
    push_within(m, pos, row, reach * reach * closest, cols);
    return closest;
}

// aggregate.rs:80-96
inline void smoothing_row(const BlockCsr<double> &m, const std::vector<Vec3> &pos, double reach,
                          const std::vector<double> &closest, size_t row, std::vector<u32> &cols) {
    double scale = closest[row];
    for (size_t at = m.starts[row]; at < m.starts[row + 1]; ++at) {
        size_t c = m.columns[at];
        if (c != row) scale = std::max(scale, closest[c]);
    }
    push_within(m, pos, row, reach * reach * scale, cols);
}

// aggregate.rs:99-178
struct Aggregation {
    std::vector<u32> aggregate;
    std::vector<char> seeded;
    u32 count = 0;
    explicit Aggregation(size_t rows) : aggregate(rows, NO_AGGREGATE), seeded(rows, 0) {}

    void seed(size_t row, const u32 *nb, size_t n) {
        if (aggregate[row] != NO_AGGREGATE || n == 0) return;
        for (size_t i = 0; i < n; ++i)
            if (aggregate[nb[i]] != NO_AGGREGATE) return;
        u32 id = count++;
        aggregate[row] = id;
        seeded[row] = 1;
        for (size_t i = 0; i < n; ++i) {
            aggregate[nb[i]] = id;
            seeded[nb[i]] = 1;
        }
    }
    void join(size_t row, const BlockCsr<double> &m, const u32 *nb, size_t n) {
        if (aggregate[row] != NO_AGGREGATE) return;
        double best = 0.0;
        u32 chosen = NO_AGGREGATE;
        for (size_t at = m.starts[row]; at < m.starts[row + 1]; ++at) {
            u32 c = m.columns[at];
            if (!std::binary_search(nb, nb + n, c) || !seeded[c]) continue;
            double norm = block_norm(m.block(at), m.block_len());
            if (norm > best) {
                best = norm;
                chosen = aggregate[c];
            }
        }
        if (chosen != NO_AGGREGATE) aggregate[row] = chosen;
    }
    void sweep(size_t row, const u32 *nb, size_t n) {
        if (aggregate[row] != NO_AGGREGATE) return;
        u32 id = count++;
        aggregate[row] = id;
        for (size_t i = 0; i < n; ++i)
            if (aggregate[nb[i]] == NO_AGGREGATE) aggregate[nb[i]] = id;
    }
};

// ---------------------------------------------------------------------------------------------
// Transfer: transfer.rs
// ---------------------------------------------------------------------------------------------

inline double column_norm(const std::vector<double> &local, size_t rows, size_t column) {
    double sum = 0.0;
    for (size_t r = 0; r < rows; ++r) {
        double v = local[r * MODES + column];
        sum += v * v;
    }
    return std::sqrt(sum);
}

// transfer.rs:19-68
inline void orthonormalize_aggregate(const u32 *members, size_t count, size_t row_block,
                                     const std::vector<double> &near_null, std::vector<double> &blocks,
                                     double coarse[36]) {
    size_t rows = count * row_block;
    std::vector<double> local(rows * MODES, 0.0);
    for (size_t m = 0; m < count; ++m) {
        const double *source = &near_null[(size_t)members[m] * row_block * MODES];
        for (size_t p = 0; p < row_block; ++p)
            for (size_t c = 0; c < MODES; ++c) local[(m * row_block + p) * MODES + c] = source[p * MODES + c];
    }
    for (int i = 0; i < 36; ++i) coarse[i] = 0.0;
    for (size_t column = 0; column < MODES; ++column) {
        double original = column_norm(local, rows, column);
        for (int pass = 0; pass < 2; ++pass)
            for (size_t previous = 0; previous < column; ++previous) {
                double projection = 0.0;
                for (size_t r = 0; r < rows; ++r) projection += local[r * MODES + previous] * local[r * MODES + column];
                coarse[previous * MODES + column] += projection;
                for (size_t r = 0; r < rows; ++r) {
                    double value = local[r * MODES + previous];
                    local[r * MODES + column] -= projection * value;
                }
            }
        double remainder = column_norm(local, rows, column);
        if (original == 0.0 || remainder <= RANK_TOLERANCE * original) {
            for (size_t r = 0; r < rows; ++r) local[r * MODES + column] = 0.0;
            coarse[column * MODES + column] = 0.0;
        } else {
            for (size_t r = 0; r < rows; ++r) local[r * MODES + column] /= remainder;
            coarse[column * MODES + column] = remainder;
        }
    }
    blocks.insert(blocks.end(), local.begin(), local.end());
}

// transfer.rs:108-152
inline void smooth_row(const BlockCsr<double> &matrix, const BlockCsr<double> &tentative, const u32 *nb, size_t nn,
                       const std::vector<double> &inverse_diagonal, double omega, size_t row, Accumulator &acc,
                       BlockCsr<double> &output) {
    size_t rb = matrix.row_block, cb = tentative.col_block, mid = matrix.col_block;
    acc.begin();
    size_t t_at = tentative.starts[row];
    u32 t_column = tentative.columns[t_at];
    acc.entry(t_column);
    for (size_t at = matrix.starts[row]; at < matrix.starts[row + 1]; ++at) {
        u32 column = matrix.columns[at];
        size_t source = ((size_t)column == row || std::binary_search(nb, nb + nn, column)) ? (size_t)column : row;
        for (size_t inner = tentative.starts[source]; inner < tentative.starts[source + 1]; ++inner) {
            double *out = acc.entry(tentative.columns[inner]);
            add_product(out, matrix.block(at), rb, mid, tentative.block(inner), cb);
        }
    }
    const double *dinv = &inverse_diagonal[row * rb * rb];
    std::vector<double> scaled(rb * cb);
    const double *t_block = tentative.block(t_at);
    acc.drain_sorted([&](u32 column, const double *block) {
        std::fill(scaled.begin(), scaled.end(), 0.0);
        add_product(scaled.data(), dinv, rb, rb, block, cb);
        output.columns.push_back(column);
        for (size_t at = 0; at < rb * cb; ++at) {
            double t_value = (column == t_column) ? t_block[at] : 0.0;
            output.values.push_back(t_value - omega * scaled[at]);
        }
    });
    output.starts.push_back(output.columns.size());
}

// transfer.rs:155-169 (with accumulate_product_row, transfer.rs:81-95)
inline void product_row(const BlockCsr<double> &matrix, const BlockCsr<double> &prolongator, size_t row,
                        Accumulator &acc, BlockCsr<double> &output) {
    size_t rb = matrix.row_block, mid = matrix.col_block, cb = prolongator.col_block;
    acc.begin();
    for (size_t at = matrix.starts[row]; at < matrix.starts[row + 1]; ++at) {
        size_t column = matrix.columns[at];
        for (size_t inner = prolongator.starts[column]; inner < prolongator.starts[column + 1]; ++inner) {
            double *out = acc.entry(prolongator.columns[inner]);
            add_product(out, matrix.block(at), rb, mid, prolongator.block(inner), cb);
        }
    }
    acc.drain_sorted([&](u32 column, const double *block) {
        output.columns.push_back(column);
        output.values.insert(output.values.end(), block, block + acc.block_len);
    });
    output.starts.push_back(output.columns.size());
}

// transfer.rs:173-194
inline void prolongator_columns(const BlockCsr<double> &p, size_t coarse_rows, std::vector<size_t> &starts,
                                std::vector<std::pair<u32, size_t>> &entries) {
    starts.assign(coarse_rows + 1, 0);
    for (u32 c : p.columns) starts[(size_t)c + 1] += 1;
    for (size_t i = 0; i < coarse_rows; ++i) starts[i + 1] += starts[i];
    std::vector<size_t> filled = starts;
    entries.assign(p.columns.size(), {0, 0});
    for (size_t row = 0; row < p.rows; ++row)
        for (size_t at = p.starts[row]; at < p.starts[row + 1]; ++at) {
            size_t c = p.columns[at];
            entries[filled[c]++] = {(u32)row, at};
        }
}

// transfer.rs:204-238
inline void galerkin_row(const BlockCsr<double> &p, const std::vector<size_t> &starts,
                         const std::vector<std::pair<u32, size_t>> &entries, const BlockCsr<double> &ap,
                         size_t coarse_row, Accumulator &acc, BlockCsr<double> &output) {
    size_t fine = p.row_block, coarse = p.col_block;
    acc.begin();
    std::vector<double> transposed(fine * coarse, 0.0);
    for (size_t e = starts[coarse_row]; e < starts[coarse_row + 1]; ++e) {
        u32 fine_row = entries[e].first;
        const double *block = p.block(entries[e].second);
        for (size_t a = 0; a < fine; ++a)
            for (size_t q = 0; q < coarse; ++q) transposed[q * fine + a] = block[a * coarse + q];
        for (size_t inner = ap.starts[fine_row]; inner < ap.starts[fine_row + 1]; ++inner) {
            double *out = acc.entry(ap.columns[inner]);
            add_product(out, transposed.data(), coarse, fine, ap.block(inner), coarse);
        }
    }
    acc.drain_sorted([&](u32 column, const double *block) {
        output.columns.push_back(column);
        output.values.insert(output.values.end(), block, block + acc.block_len);
    });
    output.starts.push_back(output.columns.size());
}

// ---------------------------------------------------------------------------------------------
// Levels and setup: setup.rs, setup_aggregation.rs (unsliced)
// ---------------------------------------------------------------------------------------------

template <typename T>
struct Level {
    BlockCsr<T> a;
    std::vector<T> inverse_diagonal;
    double lambda_min = 0, lambda_max = 0;
    bool has_prolongator = false;
    BlockCsr<T> prolongator;
    bool has_coarse = false;
    DenseCholesky coarse;
    std::vector<u32> aggregates;

    template <typename U>
    Level<U> into_precision() const {
        Level<U> l;
        l.a = a.template into_precision<U>();
        l.inverse_diagonal.resize(inverse_diagonal.size());
        for (size_t i = 0; i < inverse_diagonal.size(); ++i) l.inverse_diagonal[i] = (U)inverse_diagonal[i];
        l.lambda_min = lambda_min;
        l.lambda_max = lambda_max;
        l.has_prolongator = has_prolongator;
        if (has_prolongator) l.prolongator = prolongator.template into_precision<U>();
        l.has_coarse = has_coarse;
        l.coarse = coarse;
        l.aggregates = aggregates;
        return l;
    }
};

// setup.rs:81-695, run without slicing. Returns the f64 levels.
inline std::vector<Level<double>> build_hierarchy(const CsrMatrix &matrix, const std::vector<double> &near_null_in,
                                                  const std::vector<Vec3> &positions_in, const MultigridOptions &opt) {
    if (near_null_in.size() != MODES * matrix.size) throw std::runtime_error("near-null length");
    BlockCsr<double> current = block_from_nodal(matrix);
    if (positions_in.size() != current.rows) throw std::runtime_error("position length");
    std::vector<Vec3> positions = positions_in;
    std::vector<double> near_null = near_null_in;
    std::vector<Level<double>> levels;

    for (;;) {
        const int level_index = (int)levels.size();
        const size_t rb = current.row_block;
        // Diagonal: setup.rs:269-321
        std::vector<double> inverse_diagonal(current.rows * rb * rb, 0.0);
        for (size_t row = 0; row < current.rows; ++row) {
            auto b = current.columns.begin() + current.starts[row], e = current.columns.begin() + current.starts[row + 1];
            auto it = std::lower_bound(b, e, (u32)row);
            if (it == e || *it != (u32)row) throw SetupError(level_index, row, 0.0);
            size_t at = it - current.columns.begin();
            double pivot = 0;
            if (!invert_spd_block(current.block(at), rb, &inverse_diagonal[row * rb * rb], pivot))
                throw SetupError(level_index, row, pivot);
        }
        bool is_coarsest = current.dofs() <= opt.coarse_dofs || (size_t)level_index + 1 >= opt.max_levels;

        double reach = opt.neighbour_reach;
        size_t retries = 0;
        std::vector<size_t> strong_starts;
        std::vector<u32> strong_columns;
        std::vector<double> closest;
        Aggregation aggregation(0);
        std::vector<size_t> member_starts;
        std::vector<u32> members;
        std::vector<Vec3> coarse_positions;
        if (!is_coarsest) {
            for (;;) {
                // Strength: setup_aggregation.rs:12-43
                strong_starts.assign(1, 0);
                strong_columns.clear();
                closest.clear();
                for (size_t row = 0; row < current.rows; ++row) {
                    closest.push_back(strong_row(current, positions, reach, row, strong_columns));
                    strong_starts.push_back(strong_columns.size());
                }
                aggregation = Aggregation(current.rows);
                // Seed, Join, Sweep: setup_aggregation.rs:86-162
                for (size_t row = 0; row < current.rows; ++row)
                    aggregation.seed(row, strong_columns.data() + strong_starts[row], strong_starts[row + 1] - strong_starts[row]);
                for (size_t row = 0; row < current.rows; ++row)
                    aggregation.join(row, current, strong_columns.data() + strong_starts[row], strong_starts[row + 1] - strong_starts[row]);
                for (size_t row = 0; row < current.rows; ++row)
                    aggregation.sweep(row, strong_columns.data() + strong_starts[row], strong_starts[row + 1] - strong_starts[row]);
                // Group: setup_aggregation.rs:165-228
                size_t aggregates = aggregation.count;
                double coarse_values = (double)(aggregates * MODES);
                bool coarsens = coarse_values <= MAX_COARSE_FRACTION * (double)current.dofs();
                if (!coarsens && retries < MAX_RETRIES) {
                    retries += 1;
                    reach *= RETRY_WIDENING;
                    continue;
                }
                if (!coarsens || aggregates == 0) is_coarsest = true;
                break;
            }
        }
        double lambda_max = 0.0;
        auto spectral = [&]() {
            // setup.rs:390-449
            size_t dofs = current.dofs();
            std::vector<double> v(dofs), s1(dofs, 0.0), s2(dofs, 0.0);
            for (size_t i = 0; i < dofs; ++i) v[i] = start_value(i);
            double n0 = norm2(v.data(), dofs);
            for (size_t i = 0; i < dofs; ++i) v[i] /= n0;
            for (size_t it = 0; it < POWER_ITERATIONS; ++it) {
                current.multiply(v.data(), s1.data());
                apply_block_diagonal(inverse_diagonal.data(), current.row_block, s1.data(), s2.data(), dofs);
                double norm = norm2(s2.data(), dofs);
                lambda_max = norm;
                for (size_t i = 0; i < dofs; ++i) v[i] = s2[i] / norm;
            }
            lambda_max *= SPECTRAL_MARGIN;
        };
        if (is_coarsest) {
            spectral();
            Level<double> l;
            l.lambda_min = lambda_max / opt.smoother_range;
            l.lambda_max = lambda_max;
            if (current.dofs() <= DENSE_LIMIT) {
                // setup.rs:451-504
                size_t n = current.dofs();
                size_t cb = current.col_block;
                l.coarse.n = n;
                l.coarse.factor.assign(n * n, 0.0);
                for (size_t row = 0; row < current.rows; ++row)
                    for (size_t at = current.starts[row]; at < current.starts[row + 1]; ++at) {
                        size_t column = current.columns[at];
                        const double *b = current.block(at);
                        for (size_t p = 0; p < rb; ++p)
                            for (size_t q = 0; q < cb; ++q) l.coarse.factor[(row * rb + p) * n + column * cb + q] = b[p * cb + q];
                    }
                for (;;) {
                    size_t col = 0;
                    double pivot = 0;
                    int r = l.coarse.factor_column(col, pivot);
                    if (r < 0) throw SetupError(level_index, col, pivot);
                    if (r == 1) break;
                }
                l.has_coarse = true;
            }
            l.a = std::move(current);
            l.inverse_diagonal = std::move(inverse_diagonal);
            levels.push_back(std::move(l));
            break;
        }
        // Members and centroids: setup_aggregation.rs:186-227
        size_t aggregates = aggregation.count;
        member_starts.assign(aggregates + 1, 0);
        for (u32 a : aggregation.aggregate) member_starts[(size_t)a + 1] += 1;
        for (size_t i = 0; i < aggregates; ++i) member_starts[i + 1] += member_starts[i];
        {
            std::vector<size_t> filled = member_starts;
            members.assign(current.rows, 0);
            for (size_t row = 0; row < current.rows; ++row) members[filled[aggregation.aggregate[row]]++] = (u32)row;
        }
        coarse_positions.assign(aggregates, Vec3{{0, 0, 0}});
        for (size_t a = 0; a < aggregates; ++a) {
            for (size_t m = member_starts[a]; m < member_starts[a + 1]; ++m) {
                const Vec3 &pt = positions[members[m]];
                for (int axis = 0; axis < 3; ++axis) coarse_positions[a][axis] += pt[axis];
            }
            double count = (double)(member_starts[a + 1] - member_starts[a]);
            for (int axis = 0; axis < 3; ++axis) coarse_positions[a][axis] /= count;
        }
        // Tentative: setup.rs:327-386
        const size_t tblock = rb * MODES;
        std::vector<double> tentative_values;
        tentative_values.reserve(current.rows * tblock);
        std::vector<size_t> tentative_positions(current.rows, 0);
        std::vector<double> coarse_null;
        coarse_null.reserve(aggregates * 36);
        for (size_t a = 0; a < aggregates; ++a) {
            size_t start = tentative_values.size();
            double coarse[36];
            size_t count = member_starts[a + 1] - member_starts[a];
            orthonormalize_aggregate(&members[member_starts[a]], count, rb, near_null, tentative_values, coarse);
            for (size_t i = 0; i < count; ++i) tentative_positions[members[member_starts[a] + i]] = start + i * tblock;
            coarse_null.insert(coarse_null.end(), coarse, coarse + 36);
        }
        auto tentative = BlockCsr<double>::with_shape(current.rows, rb, MODES);
        for (size_t row = 0; row < current.rows; ++row) {
            tentative.columns.push_back(aggregation.aggregate[row]);
            size_t at = tentative_positions[row];
            tentative.values.insert(tentative.values.end(), &tentative_values[at], &tentative_values[at] + tblock);
            tentative.starts.push_back(tentative.columns.size());
        }
        spectral();
        // Filter: setup_aggregation.rs:48-84 (not on the finest level)
        std::vector<size_t> filter_starts;
        std::vector<u32> filter_columns;
        if (level_index != 0) {
            filter_starts.assign(1, 0);
            for (size_t row = 0; row < current.rows; ++row) {
                smoothing_row(current, positions, reach, closest, row, filter_columns);
                filter_starts.push_back(filter_columns.size());
            }
        }
        // Smooth: setup.rs:527-573
        double omega = 4.0 / (3.0 * lambda_max);
        Accumulator acc(aggregates, rb * MODES);
        auto prolongator = BlockCsr<double>::with_shape(current.rows, rb, MODES);
        for (size_t row = 0; row < current.rows; ++row) {
            const u32 *nb;
            size_t nn;
            if (filter_starts.empty()) {
                nb = strong_columns.data() + strong_starts[row];
                nn = strong_starts[row + 1] - strong_starts[row];
            } else {
                nb = filter_columns.data() + filter_starts[row];
                nn = filter_starts[row + 1] - filter_starts[row];
            }
            smooth_row(current, tentative, nb, nn, inverse_diagonal, omega, row, acc, prolongator);
        }
        // Product: setup.rs:575-612
        auto a_times_p = BlockCsr<double>::with_shape(current.rows, rb, MODES);
        for (size_t row = 0; row < current.rows; ++row) product_row(current, prolongator, row, acc, a_times_p);
        std::vector<size_t> column_starts;
        std::vector<std::pair<u32, size_t>> column_entries;
        prolongator_columns(prolongator, aggregates, column_starts, column_entries);
        // Galerkin: setup.rs:614-647
        auto coarse = BlockCsr<double>::with_shape(aggregates, MODES, MODES);
        Accumulator acc2(aggregates, MODES * MODES);
        for (size_t c = 0; c < aggregates; ++c) galerkin_row(prolongator, column_starts, column_entries, a_times_p, c, acc2, coarse);
        // finish_level: setup.rs:650-695
        size_t dim = coarse.row_block;
        for (size_t row = 0; row < coarse.rows; ++row)
            for (size_t at = coarse.starts[row]; at < coarse.starts[row + 1]; ++at) {
                if ((size_t)coarse.columns[at] != row) continue;
                size_t block_at = at * dim * dim;
                for (size_t d = 0; d < dim; ++d) {
                    size_t idx = block_at + d * dim + d;
                    if (coarse.values[idx] == 0.0) coarse.values[idx] = 1.0;
                }
            }
        Level<double> l;
        l.lambda_min = lambda_max / opt.smoother_range;
        l.lambda_max = lambda_max;
        l.a = std::move(current);
        l.inverse_diagonal = std::move(inverse_diagonal);
        l.has_prolongator = true;
        l.prolongator = std::move(prolongator);
        l.aggregates = aggregation.aggregate;
        levels.push_back(std::move(l));
        current = std::move(coarse);
        near_null = std::move(coarse_null);
        positions = std::move(coarse_positions);
    }
    return levels;
}

// ---------------------------------------------------------------------------------------------
// The cycle: cycle.rs:461-674, generic over storage S and vector V
// ---------------------------------------------------------------------------------------------

template <typename V>
struct LevelWork {
    std::vector<V> residual, scaled, direction, product, post_rhs, post_x, coarse_rhs, coarse_x;
    std::vector<double> dense;
    // cycle.rs:84-106: this level's row chunks, one when it runs on one thread
    std::vector<size_t> bounds;
};

template <typename S, typename V>
inline void apply_inverse_diagonal(const Level<S> &level, size_t first, size_t last, const V *x, V *y) {
    size_t block = level.a.row_block;
    for (size_t row = first; row < last; ++row) {
        const S *d = &level.inverse_diagonal[row * block * block];
        size_t local = row - first;
        for (size_t p = 0; p < block; ++p) {
            V sum = (V)0;
            for (size_t q = 0; q < block; ++q) sum += widen<V>(d[p * block + q]) * x[local * block + q];
            y[local * block + p] = sum;
        }
    }
}

// cycle.rs:570-651, each region over the level's row chunks
template <typename S, typename V>
inline void smooth(const Level<S> &level, size_t degree, WorkerPool *pool, LevelWork<V> &w, const V *b, V *x) {
    double theta = 0.5 * (level.lambda_max + level.lambda_min);
    double delta = 0.5 * (level.lambda_max - level.lambda_min);
    double sigma = theta / delta;
    double rho = 1.0 / sigma;
    const size_t block = level.a.row_block;
    V first = (V)theta;
    for_rows(pool, w.bounds, [&](size_t r0, size_t r1) {
        const size_t i0 = r0 * block, i1 = r1 * block;
        for (size_t i = i0; i < i1; ++i) w.residual[i] = b[i];
        apply_inverse_diagonal(level, r0, r1, w.residual.data() + i0, w.scaled.data() + i0);
        for (size_t i = i0; i < i1; ++i) w.direction[i] = w.scaled[i] / first;
        for (size_t i = i0; i < i1; ++i) x[i] = w.direction[i];
    });
    for (size_t step = 1; step < degree; ++step) {
        // the product reads every row of the direction: a region of its own
        for_rows(pool, w.bounds, [&](size_t r0, size_t r1) {
            level.a.multiply_rows(r0, r1, w.direction.data(), w.product.data() + r0 * block);
        });
        double next_rho = 1.0 / (2.0 * sigma - rho);
        V keep = (V)(next_rho * rho);
        V stp = (V)(2.0 * next_rho / delta);
        for_rows(pool, w.bounds, [&](size_t r0, size_t r1) {
            const size_t i0 = r0 * block, i1 = r1 * block;
            for (size_t i = i0; i < i1; ++i) w.residual[i] -= w.product[i];
            apply_inverse_diagonal(level, r0, r1, w.residual.data() + i0, w.scaled.data() + i0);
            for (size_t i = i0; i < i1; ++i) w.direction[i] = keep * w.direction[i] + stp * w.scaled[i];
            for (size_t i = i0; i < i1; ++i) x[i] += w.direction[i];
        });
        rho = next_rho;
    }
}

// cycle.rs:548-564
template <typename S, typename V>
inline void residual(const Level<S> &level, WorkerPool *pool, const std::vector<size_t> &bounds, const V *b, const V *x, V *r) {
    const size_t block = level.a.row_block;
    for_rows(pool, bounds, [&](size_t r0, size_t r1) {
        level.a.multiply_rows(r0, r1, x, r + r0 * block);
        for (size_t i = r0 * block; i < r1 * block; ++i) r[i] = b[i] - r[i];
    });
}

// cycle.rs:461-539
template <typename S, typename V>
inline void cycle(const std::vector<Level<S>> &levels, size_t index, size_t degree, WorkerPool *pool, std::vector<LevelWork<V>> &work,
                  const V *rhs, V *x) {
    const Level<S> &level = levels[index];
    LevelWork<V> &lw = work[index];
    if (!level.has_prolongator) {
        if (level.has_coarse) {
            size_t n = level.a.dofs();
            for (size_t i = 0; i < n; ++i) lw.dense[i] = (double)rhs[i];
            level.coarse.solve(lw.dense.data());
            for (size_t i = 0; i < n; ++i) x[i] = (V)lw.dense[i];
        } else {
            smooth(level, degree, pool, lw, rhs, x);
        }
        return;
    }
    const size_t block = level.a.row_block;
    smooth(level, degree, pool, lw, rhs, x);
    residual(level, pool, lw.bounds, rhs, x, lw.residual.data());
    level.prolongator.multiply_transpose(lw.residual.data(), lw.coarse_rhs.data(), lw.coarse_rhs.size());
    std::fill(lw.coarse_x.begin(), lw.coarse_x.end(), (V)0);
    cycle(levels, index + 1, degree, pool, work, lw.coarse_rhs.data(), lw.coarse_x.data());
    for_rows(pool, lw.bounds, [&](size_t r0, size_t r1) {
        level.prolongator.multiply_rows(r0, r1, lw.coarse_x.data(), lw.product.data() + r0 * block);
        for (size_t i = r0 * block; i < r1 * block; ++i) x[i] += lw.product[i];
    });
    residual(level, pool, lw.bounds, rhs, x, lw.post_rhs.data());
    std::vector<V> post_rhs = std::move(lw.post_rhs);
    std::vector<V> post_x = std::move(lw.post_x);
    smooth(level, degree, pool, lw, post_rhs.data(), post_x.data());
    for_rows(pool, lw.bounds, [&](size_t r0, size_t r1) {
        for (size_t i = r0 * block; i < r1 * block; ++i) x[i] += post_x[i];
    });
    lw.post_rhs = std::move(post_rhs);
    lw.post_x = std::move(post_x);
}

template <typename S, typename V>
inline std::vector<LevelWork<V>> make_work(const std::vector<Level<S>> &levels, size_t threads) {
    std::vector<LevelWork<V>> work(levels.size());
    for (size_t i = 0; i < levels.size(); ++i) {
        size_t dofs = levels[i].a.dofs();
        size_t cd = (i + 1 < levels.size()) ? levels[i + 1].a.dofs() : 0;
        auto &w = work[i];
        w.residual.assign(dofs, 0);
        w.scaled.assign(dofs, 0);
        w.direction.assign(dofs, 0);
        w.product.assign(dofs, 0);
        w.post_rhs.assign(dofs, 0);
        w.post_x.assign(dofs, 0);
        w.coarse_rhs.assign(cd, 0);
        w.coarse_x.assign(cd, 0);
        if (levels[i].has_coarse) w.dense.assign(dofs, 0.0);
        const size_t chunks = levels[i].a.values.size() >= MIN_PARALLEL_VALUES ? threads : 1;
        w.bounds = balanced_bounds(levels[i].a.starts, chunks);
    }
    return work;
}

// The finished preconditioner: cycle.rs:246-458
struct Multigrid {
    Precision precision = Precision::Double;
    size_t degree = 3;
    std::vector<Level<double>> levels64;  // as built (f64), kept for inspection
    std::vector<Level<float>> levels32;   // rounded, for the f32-stored precisions
    mutable std::vector<LevelWork<double>> work64;
    mutable std::vector<LevelWork<float>> work32;
    mutable std::vector<float> rhs32, x32;
    // cycle.rs:265-308: a pool when more than one thread is asked for and some level is large
    // enough to use it
    std::unique_ptr<WorkerPool> pool;

    Multigrid(std::vector<Level<double>> levels, size_t deg, Precision p, size_t threads = 1)
        : precision(p), degree(deg), levels64(std::move(levels)) {
        bool large = false;
        for (const auto &l : levels64) large = large || l.a.values.size() >= MIN_PARALLEL_VALUES;
        if (large) pool = make_pool(threads);
        const size_t used = pool ? pool->threads() : 1;
        if (precision == Precision::Double) {
            work64 = make_work<double, double>(levels64, used);
        } else {
            for (auto &l : levels64) levels32.push_back(l.into_precision<float>());
            if (precision == Precision::SingleStorage) {
                work64 = make_work<float, double>(levels32, used);
            } else {
                work32 = make_work<float, float>(levels32, used);
                rhs32.assign(levels32[0].a.dofs(), 0.0f);
                x32.assign(levels32[0].a.dofs(), 0.0f);
            }
        }
    }

    // cycle.rs:430-446
    void apply(const double *r, double *z) const {
        switch (precision) {
            case Precision::Double:
                cycle<double, double>(levels64, 0, degree, pool.get(), work64, r, z);
                break;
            case Precision::SingleStorage:
                cycle<float, double>(levels32, 0, degree, pool.get(), work64, r, z);
                break;
            case Precision::Single: {
                size_t n = rhs32.size();
                for (size_t i = 0; i < n; ++i) rhs32[i] = (float)r[i];
                cycle<float, float>(levels32, 0, degree, pool.get(), work32, rhs32.data(), x32.data());
                for (size_t i = 0; i < n; ++i) z[i] = (double)x32[i];
                break;
            }
        }
    }

    // one smoother application on `level` from a zero guess, in the hierarchy's precision
    void smooth_level(size_t level, const double *b, double *x) const {
        switch (precision) {
            case Precision::Double:
                smooth<double, double>(levels64[level], degree, pool.get(), work64[level], b, x);
                break;
            case Precision::SingleStorage:
                smooth<float, double>(levels32[level], degree, pool.get(), work64[level], b, x);
                break;
            case Precision::Single: {
                size_t n = levels32[level].a.dofs();
                std::vector<float> bb(n), xx(n);
                for (size_t i = 0; i < n; ++i) bb[i] = (float)b[i];
                smooth<float, float>(levels32[level], degree, pool.get(), work32[level], bb.data(), xx.data());
                for (size_t i = 0; i < n; ++i) x[i] = (double)xx[i];
                break;
            }
        }
    }

    size_t threads() const { return pool ? pool->threads() : 1; }

    // Multigrid::system_multiply, cycle.rs:447-457: the conjugate gradient's A p on the pool
    // once the matrix holds enough values
    void system_multiply(const CsrMatrix &matrix, const double *x, double *y) const {
        if (pool && matrix.values.size() >= MIN_PARALLEL_VALUES) {
            const auto bounds = balanced_bounds(matrix.row_starts, pool->threads());
            for_rows(pool.get(), bounds, [&](size_t r0, size_t r1) { matrix.multiply_rows(r0, r1, x, y); });
        } else {
            matrix.multiply(x, y);
        }
    }

    // MultigridStats::of_levels, cycle.rs:676-699
    void stats(size_t &nlevels, size_t &coarse_dofs, double &oc, size_t &memory) const {
        nlevels = levels64.size();
        coarse_dofs = levels64.empty() ? 0 : levels64.back().a.dofs();
        double finest = levels64.empty() ? 1.0 : (double)levels64[0].a.values.size();
        size_t stored = 0;
        memory = 0;
        size_t tsize = precision == Precision::Double ? sizeof(double) : sizeof(float);
        for (const auto &l : levels64) {
            stored += l.a.values.size();
            memory += l.a.starts.size() * sizeof(size_t) + l.a.columns.size() * sizeof(u32) + l.a.values.size() * tsize;
            memory += l.inverse_diagonal.size() * tsize;
            if (l.has_prolongator)
                memory += l.prolongator.starts.size() * sizeof(size_t) + l.prolongator.columns.size() * sizeof(u32) +
                          l.prolongator.values.size() * tsize;
            if (l.has_coarse) memory += l.coarse.memory_bytes();
        }
        oc = (double)stored / finest;
    }
};

}  // namespace refcpu
