// The numeric phases of a multigrid level's setup on the device: the power iteration
// (crates/solidfem-solve/src/multigrid/setup.rs:390-449), the smoothed prolongator
// P = T - omega D^-1 A_f T (setup.rs:527-573, transfer.rs:108-152), A P (setup.rs:575-612,
// transfer.rs:155-169) and the Galerkin product P^T (A P) (setup.rs:614-647,
// transfer.rs:204-238).
//
// Every output value is the host phase's value to the bit. The host sums each output entry in
// a fixed order: contributions arrive row entry by row entry, and each block product adds its
// terms in ascending inner index, one rounding per add, with no fused multiply-add (the library
// is built with -ffp-contract=off). Here one wavefront owns one output row and one lane owns one
// output block: every lane walks the row's contributions in the host's order and adds the ones
// that land in its block, so each entry sees the same sequence of roundings.
//
// An output row's block columns are the sorted distinct coarse indices of its contributions.
// Coarse indices are aggregate numbers, so one pass marks them in a bitmap in LDS over the
// row's index window and a second reads the set bits in ascending order. A count launch sizes
// the rows (host prefix sum), a value launch fills them.
//
// The power iteration's norm is a serial sum over the vector (setup.rs:418-420); its chain of
// dependent additions runs on the host over a copy of the vector, the products on the device.
#include <algorithm>
#include <cmath>
#include <climits>

#include "sf_device.hpp"
#include "sf_setup_device.hpp"

namespace sf {
namespace {

constexpr int WAVE_SIZE = 64;
constexpr u32 WINDOW_WORDS = 2048;  // a row's coarse indices must fit 65,536 consecutive values
constexpr u32 MAX_SLOTS = 1024;     // blocks per output row
constexpr int MODES6 = 6;

__device__ __forceinline__ u32 wave_min_u32(u32 v) {
    for (int o = 32; o; o >>= 1) v = min(v, (u32)__shfl_xor((int)v, o));
    return v;
}
__device__ __forceinline__ u32 wave_max_u32(u32 v) {
    for (int o = 32; o; o >>= 1) v = max(v, (u32)__shfl_xor((int)v, o));
    return v;
}
// Exclusive prefix over the wave's lanes; `total` gets the sum.
__device__ __forceinline__ u32 wave_exclusive_u32(u32 v, u32 &total) {
    const int lane = threadIdx.x;
    u32 inc = v;
    for (int o = 1; o < WAVE_SIZE; o <<= 1) {
        const u32 up = (u32)__shfl_up((int)inc, o);
        if (lane >= o) inc += up;
    }
    total = (u32)__shfl((int)inc, WAVE_SIZE - 1);
    return inc - v;
}

// The sorted distinct coarse indices one row's contributions land in. `each(lane, f)` calls
// f(index) for every contribution a lane enumerates (lanes split the contributions among them).
// Leaves them in `slots[0..k)` and returns k, or UINT_MAX if the row overflows the window or
// holds more than `limit` (at most MAX_SLOTS) columns: the caller reports the overflow and the
// level's phase runs on the host.
template <typename Each>
__device__ u32 row_columns(Each each, u32 *bitmap, u32 *slots, bool keep, u32 limit) {
    const int lane = threadIdx.x;
    u32 lo = UINT_MAX, hi = 0;
    each(lane, [&](u32 id) {
        lo = min(lo, id);
        hi = max(hi, id);
    });
    lo = wave_min_u32(lo);
    hi = wave_max_u32(hi);
    if (lo > hi) return 0;
    const u32 wlo = lo >> 5, words = (hi >> 5) - wlo + 1;
    if (words > WINDOW_WORDS) return UINT_MAX;
    for (u32 w = lane; w < words; w += WAVE_SIZE) bitmap[w] = 0;
    __syncthreads();
    each(lane, [&](u32 id) { atomicOr(&bitmap[(id >> 5) - wlo], 1u << (id & 31)); });
    __syncthreads();
    u32 base = 0;
    for (u32 w0 = 0; w0 < words; w0 += WAVE_SIZE) {
        u32 bits = w0 + lane < words ? bitmap[w0 + lane] : 0u;
        u32 total;
        u32 pos = base + wave_exclusive_u32((u32)__popc(bits), total);
        if (keep) {
            while (bits) {
                const int b = __ffs(bits) - 1;
                bits &= bits - 1;
                if (pos < MAX_SLOTS) slots[pos] = ((wlo + w0 + (u32)lane) << 5) + (u32)b;
                ++pos;
            }
        }
        base += total;
    }
    __syncthreads();
    return base > limit ? UINT_MAX : base;
}

__device__ __forceinline__ bool sorted_contains(const u32 *first, int64_t n, u32 value) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (first[mid] < value)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < n && first[lo] == value;
}

// y = D^-1 (A x), one thread per block row, each sum in the host's order (block.rs:136-154,
// setup.rs:409-417).
template <int RB>
__global__ __launch_bounds__(256) void power_kernel(size_t rows, const int64_t *__restrict__ starts, const u32 *__restrict__ cols,
                                                    const double *__restrict__ vals, const double *__restrict__ dinv,
                                                    const double *__restrict__ x, double *__restrict__ y) {
    const size_t row = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    double out[RB];
#pragma unroll
    for (int p = 0; p < RB; ++p) out[p] = 0.0;
    for (int64_t at = starts[row]; at < starts[row + 1]; ++at) {
        const double *b = vals + (size_t)at * RB * RB;
        const double *xc = x + (size_t)cols[at] * RB;
        double xv[RB];
#pragma unroll
        for (int q = 0; q < RB; ++q) xv[q] = xc[q];
#pragma unroll
        for (int p = 0; p < RB; ++p) {
            double s = out[p];
#pragma unroll
            for (int q = 0; q < RB; ++q) s += b[p * RB + q] * xv[q];
            out[p] = s;
        }
    }
    const double *d = dinv + row * RB * RB;
#pragma unroll
    for (int p = 0; p < RB; ++p) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < RB; ++q) s += d[p * RB + q] * out[q];
        y[row * RB + p] = s;
    }
}

__global__ void divide_kernel(size_t n, const double *__restrict__ y, double norm, double *__restrict__ x) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = y[i] / norm;
}

// The blocks at `slots`, in that order.
__global__ void gather_blocks_kernel(size_t count, int len, const int64_t *__restrict__ slots, const double *__restrict__ values,
                                     double *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count * (size_t)len) return;
    const size_t b = i / (size_t)len, k = i % (size_t)len;
    out[i] = values[(size_t)slots[b] * (size_t)len + k];
}

// Per block, the sum of its squared values in storage order from zero (the host's order).
__global__ void block_squares_kernel(size_t blocks, int len, const double *__restrict__ values, double *__restrict__ out) {
    const size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= blocks) return;
    const double *v = values + b * (size_t)len;
    double s = 0.0;
    for (int k = 0; k < len; ++k) s += v[k] * v[k];
    out[b] = s;
}

struct SetupRows {
    const int64_t *starts;
    const u32 *cols;
    const double *vals;
};

// P = T - omega D^-1 (A_f T): A's block (row, col) carries T's row col when col is the row or
// one of its strong (fine level) or filtered (coarse levels) neighbours, else T's own row
// (transfer.rs:108-152, aggregate.rs:80-96). COUNT launches write the row lengths.
template <int RB, bool COUNT>
__global__ __launch_bounds__(64) void smooth_kernel(size_t rows, SetupRows a, const u32 *__restrict__ aggregate,
                                                    const int64_t *__restrict__ ns, const u32 *__restrict__ nc,
                                                    const double *__restrict__ tentative, const double *__restrict__ dinv,
                                                    double omega, u32 *__restrict__ counts, const int64_t *__restrict__ out_starts,
                                                    u32 *__restrict__ out_cols, double *__restrict__ out_vals, u32 *overflow, u32 slot_limit) {
    __shared__ u32 bitmap[WINDOW_WORDS];
    __shared__ u32 slots[MAX_SLOTS];
    const size_t row = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t first = a.starts[row], last = a.starts[row + 1];
    const u32 *nb = nc + ns[row];
    const int64_t nn = ns[row + 1] - ns[row];
    auto source = [&](int64_t at) -> u32 {
        const u32 col = a.cols[at];
        return (col == (u32)row || sorted_contains(nb, nn, col)) ? col : (u32)row;
    };
    const u32 k = row_columns(
        [&](int l, auto f) {
            for (int64_t at = first + l; at < last; at += WAVE_SIZE) f(aggregate[source(at)]);
        },
        bitmap, slots, !COUNT, slot_limit);
    if (k == UINT_MAX) {
        if (lane == 0) atomicOr(overflow, 1u);
        return;
    }
    if (COUNT) {
        if (lane == 0) counts[row] = k;
        return;
    }
    const u32 t_col = aggregate[row];
    const double *d = dinv + row * RB * RB;
    const double *t_row = tentative + row * RB * MODES6;
    const int64_t base = out_starts[row];
    for (u32 g = 0; g < k; g += WAVE_SIZE) {
        const u32 s = g + (u32)lane;
        const u32 mine = s < k ? slots[s] : UINT_MAX;
        double acc[RB * MODES6];
#pragma unroll
        for (int i = 0; i < RB * MODES6; ++i) acc[i] = 0.0;
        for (int64_t at = first; at < last; ++at) {
            const u32 src = source(at);
            if (aggregate[src] != mine) continue;
            const double *b = a.vals + (size_t)at * RB * RB;
            const double *t = tentative + (size_t)src * RB * MODES6;
#pragma unroll
            for (int r = 0; r < RB; ++r)
#pragma unroll
                for (int mid = 0; mid < RB; ++mid) {
                    const double v = b[r * RB + mid];
#pragma unroll
                    for (int q = 0; q < MODES6; ++q) acc[r * MODES6 + q] += v * t[mid * MODES6 + q];
                }
        }
        if (s >= k) continue;
        out_cols[base + s] = mine;
        double *out = out_vals + (size_t)(base + s) * RB * MODES6;
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int q = 0; q < MODES6; ++q) {
                double scaled = 0.0;
#pragma unroll
                for (int mid = 0; mid < RB; ++mid) scaled += d[r * RB + mid] * acc[mid * MODES6 + q];
                out[r * MODES6 + q] = (mine == t_col ? t_row[r * MODES6 + q] : 0.0) - omega * scaled;
            }
    }
}

// A P by fine rows: A's blocks in order, each times the prolongator row of its column, the
// prolongator row's blocks in order (transfer.rs:155-169).
template <int RB, bool COUNT>
__global__ __launch_bounds__(64) void product_kernel(size_t rows, SetupRows a, SetupRows p, u32 *__restrict__ counts,
                                                     const int64_t *__restrict__ out_starts, u32 *__restrict__ out_cols,
                                                     double *__restrict__ out_vals, u32 *overflow, u32 slot_limit) {
    __shared__ u32 bitmap[WINDOW_WORDS];
    __shared__ u32 slots[MAX_SLOTS];
    const size_t row = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t first = a.starts[row], last = a.starts[row + 1];
    const u32 k = row_columns(
        [&](int l, auto f) {
            for (int64_t at = first + l; at < last; at += WAVE_SIZE) {
                const u32 col = a.cols[at];
                for (int64_t in = p.starts[col]; in < p.starts[col + 1]; ++in) f(p.cols[in]);
            }
        },
        bitmap, slots, !COUNT, slot_limit);
    if (k == UINT_MAX) {
        if (lane == 0) atomicOr(overflow, 1u);
        return;
    }
    if (COUNT) {
        if (lane == 0) counts[row] = k;
        return;
    }
    const int64_t base = out_starts[row];
    for (u32 g = 0; g < k; g += WAVE_SIZE) {
        const u32 s = g + (u32)lane;
        const u32 mine = s < k ? slots[s] : UINT_MAX;
        double acc[RB * MODES6];
#pragma unroll
        for (int i = 0; i < RB * MODES6; ++i) acc[i] = 0.0;
        for (int64_t at = first; at < last; ++at) {
            const u32 col = a.cols[at];
            const double *b = a.vals + (size_t)at * RB * RB;
            for (int64_t in = p.starts[col]; in < p.starts[col + 1]; ++in) {
                if (p.cols[in] != mine) continue;
                const double *pb = p.vals + (size_t)in * RB * MODES6;
#pragma unroll
                for (int r = 0; r < RB; ++r)
#pragma unroll
                    for (int mid = 0; mid < RB; ++mid) {
                        const double v = b[r * RB + mid];
#pragma unroll
                        for (int q = 0; q < MODES6; ++q) acc[r * MODES6 + q] += v * pb[mid * MODES6 + q];
                    }
            }
        }
        if (s >= k) continue;
        out_cols[base + s] = mine;
        double *out = out_vals + (size_t)(base + s) * RB * MODES6;
#pragma unroll
        for (int i = 0; i < RB * MODES6; ++i) out[i] = acc[i];
    }
}

// P^T (A P) by coarse rows: the prolongator's entries of column c by ascending fine row, each
// transposed block times that fine row of A P, its blocks in order (transfer.rs:204-238).
template <int RB, bool COUNT>
__global__ __launch_bounds__(64) void galerkin_kernel(size_t count, const int64_t *__restrict__ pcol_starts,
                                                      const u32 *__restrict__ pcol_rows, const int64_t *__restrict__ pcol_at,
                                                      const double *__restrict__ pvals, SetupRows ap, u32 *__restrict__ counts,
                                                      const int64_t *__restrict__ out_starts, u32 *__restrict__ out_cols,
                                                      double *__restrict__ out_vals, u32 *overflow, u32 slot_limit) {
    __shared__ u32 bitmap[WINDOW_WORDS];
    __shared__ u32 slots[MAX_SLOTS];
    const size_t c = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t first = pcol_starts[c], last = pcol_starts[c + 1];
    const u32 k = row_columns(
        [&](int l, auto f) {
            for (int64_t e = first + l; e < last; e += WAVE_SIZE) {
                const u32 fine = pcol_rows[e];
                for (int64_t in = ap.starts[fine]; in < ap.starts[fine + 1]; ++in) f(ap.cols[in]);
            }
        },
        bitmap, slots, !COUNT, slot_limit);
    if (k == UINT_MAX) {
        if (lane == 0) atomicOr(overflow, 1u);
        return;
    }
    if (COUNT) {
        if (lane == 0) counts[c] = k;
        return;
    }
    const int64_t base = out_starts[c];
    for (u32 g = 0; g < k; g += WAVE_SIZE) {
        const u32 s = g + (u32)lane;
        const u32 mine = s < k ? slots[s] : UINT_MAX;
        double acc[MODES6 * MODES6];
#pragma unroll
        for (int i = 0; i < MODES6 * MODES6; ++i) acc[i] = 0.0;
        for (int64_t e = first; e < last; ++e) {
            const u32 fine = pcol_rows[e];
            const double *pb = pvals + (size_t)pcol_at[e] * RB * MODES6;
            for (int64_t in = ap.starts[fine]; in < ap.starts[fine + 1]; ++in) {
                if (ap.cols[in] != mine) continue;
                const double *apb = ap.vals + (size_t)in * RB * MODES6;
#pragma unroll
                for (int r = 0; r < MODES6; ++r)
#pragma unroll
                    for (int mid = 0; mid < RB; ++mid) {
                        const double v = pb[mid * MODES6 + r];
#pragma unroll
                        for (int q = 0; q < MODES6; ++q) acc[r * MODES6 + q] += v * apb[mid * MODES6 + q];
                    }
            }
        }
        if (s >= k) continue;
        // setup.rs:650-695: a unit diagonal where no aggregate direction reaches (the host applies
        // the same rule to the finished matrix; it leaves these values as they are)
        if (mine == (u32)c)
#pragma unroll
            for (int d = 0; d < MODES6; ++d)
                if (acc[d * MODES6 + d] == 0.0) acc[d * MODES6 + d] = 1.0;
        out_cols[base + s] = mine;
        double *out = out_vals + (size_t)(base + s) * MODES6 * MODES6;
#pragma unroll
        for (int i = 0; i < MODES6 * MODES6; ++i) out[i] = acc[i];
    }
}

size_t grid_of(size_t n, size_t threads) { return (n + threads - 1) / threads; }

// Host prefix sum of per-row counts into row starts.
std::vector<int64_t> starts_of(const std::vector<u32> &counts) {
    std::vector<int64_t> starts(counts.size() + 1, 0);
    for (size_t i = 0; i < counts.size(); ++i) starts[i + 1] = starts[i] + counts[i];
    return starts;
}

// ---- the seed pass of aggregation, in rounds ----
//
// The serial pass (aggregate.rs:119-136) visits the rows in index order, and a row seeds when it
// and its strong neighbours (its closed neighbourhood) are all still free. That is the greedy
// independent set, in index order, of the graph in which two rows meet when their closed
// neighbourhoods share a row. The rounds below build that same set: in each, every undecided row
// claims the rows of its closed neighbourhood with its index (the least claim wins); a row that
// wins all of its own seeds, since no undecided row of lower index shares a row with it, and the
// serial pass would find its neighbourhood free; then every undecided row whose neighbourhood
// meets a seed's is out. The least undecided row seeds in every round, so the rounds end.
constexpr uint8_t SEED_UNDECIDED = 0, SEED_IN = 1, SEED_OUT = 2, SEED_NEVER = 3;

__global__ void seed_claim_kernel(u32 rows, const u32 *__restrict__ starts, const u32 *__restrict__ cols,
                                  const uint8_t *__restrict__ state, u32 *__restrict__ bid) {
    const u32 r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows || state[r] != SEED_UNDECIDED) return;
    atomicMin(&bid[r], r);
    for (u32 k = starts[r]; k < starts[r + 1]; ++k) atomicMin(&bid[cols[k]], r);
}

__global__ void seed_decide_kernel(u32 rows, const u32 *__restrict__ starts, const u32 *__restrict__ cols,
                                   uint8_t *__restrict__ state, const u32 *__restrict__ bid, u32 *__restrict__ owner) {
    const u32 r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows || state[r] != SEED_UNDECIDED || bid[r] != r) return;
    for (u32 k = starts[r]; k < starts[r + 1]; ++k)
        if (bid[cols[k]] != r) return;
    // two seeds of one round never share a row: each won all of its own
    state[r] = SEED_IN;
    owner[r] = r;
    for (u32 k = starts[r]; k < starts[r + 1]; ++k) owner[cols[k]] = r;
}

__global__ void seed_settle_kernel(u32 rows, const u32 *__restrict__ starts, const u32 *__restrict__ cols,
                                   uint8_t *__restrict__ state, const u32 *__restrict__ owner, u32 *__restrict__ bid,
                                   u32 *__restrict__ undecided) {
    const u32 r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    bid[r] = UINT_MAX;  // the next round's claims start from none
    if (state[r] != SEED_UNDECIDED) return;
    bool out = owner[r] != UINT_MAX;
    for (u32 k = starts[r]; k < starts[r + 1] && !out; ++k) out = owner[cols[k]] != UINT_MAX;
    if (out) state[r] = SEED_OUT;
    else if (undecided) *undecided = 1u;
}

// The dense Cholesky factor (block.rs:327-348) in place, right-looking: at step k, column k is
// finished (its entries below the diagonal divided by the root of the diagonal), then every entry
// of the trailing lower triangle takes its term a[i][k] a[j][k]. An entry thus receives its
// subtractions one per step, in ascending k, and its division last: the operations, in the order,
// of the host's column loop, whose bits it has. The diagonal a[k][k] is read by step k's division
// and by nothing after, so the roots are stored at the end (the value under each root is kept for
// the host's refusal of a matrix that is not positive definite).
__global__ void cholesky_finish_column_kernel(u32 n, u32 k, double *__restrict__ a, double *__restrict__ pivots) {
    const u32 i = k + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = a[(size_t)k * n + k];
    if (i == k) pivots[k] = s;
    else a[(size_t)i * n + k] = a[(size_t)i * n + k] / sqrt(s);
}

__global__ void cholesky_update_kernel(u32 n, u32 k, double *__restrict__ a) {
    const u32 j = k + 1 + blockIdx.x * blockDim.x + threadIdx.x, i = k + 1 + blockIdx.y;
    if (i >= n || j > i) return;
    a[(size_t)i * n + j] -= a[(size_t)i * n + k] * a[(size_t)j * n + k];
}

__global__ void cholesky_diagonal_kernel(u32 n, const double *__restrict__ pivots, double *__restrict__ a) {
    const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) a[(size_t)j * n + j] = sqrt(pivots[j]);
}

__global__ void fill_u32_kernel(size_t n, u32 value, u32 *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = value;
}

}  // namespace

std::atomic<u32> &setup_slot_limit() {
    // SOLIDFEM_SETUP_SLOT_LIMIT gives the starting value (read once), sf_setup_slot_limit changes
    // it; 0 means the kernels' own limit
    static std::atomic<u32> limit{[] {
        const char *e = std::getenv("SOLIDFEM_SETUP_SLOT_LIMIT");
        return e ? (u32)std::max<long>(0, std::atol(e)) : 0u;
    }()};
    return limit;
}

DeviceSetup::DeviceSetup(std::shared_ptr<Context> ctx, size_t min_rows, const double *level0_values, bool aggregate_on_device)
    : ctx_(std::move(ctx)), min_rows_(min_rows), level0_values_(level0_values), aggregate_on_device_(aggregate_on_device) {
    // fewer output blocks per row than the kernels hold, so that rows above it take the host
    // path (the fallback's own bits)
    if (const u32 limit = setup_slot_limit().load(std::memory_order_relaxed)) slot_limit_ = std::min<u32>(MAX_SLOTS, limit);
}

DeviceSetup::~DeviceSetup() {
    if (ctx_) (void)hipStreamSynchronize(ctx_->stream);
}

bool DeviceSetup::takes(int level, const HostBsr<double> &a) const {
    return a.rows >= min_rows_ && (a.rb == 3 || a.rb == 6) && a.cb == a.rb && a.rows < (size_t)UINT_MAX;
}

void DeviceSetup::begin(int level, const HostBsr<double> &a, const std::vector<double> &inverse_diagonal, const double *device_values) {
    SF_HIP(hipSetDevice(ctx_->device));  // may run on a thread of its own
    hipStream_t s = ctx_->stream;
    level_ = level;
    rb_ = (int)a.rb;
    rows_ = a.rows;
    a_starts_.upload(a.starts, s);
    a_cols_.upload(a.columns, s);
    if (device_values) {
        a_vals_.release();
        a_vals_ptr_ = device_values;
    } else if (level == 0 && level0_values_) {
        a_vals_.release();
        a_vals_ptr_ = level0_values_;
    } else {
        a_vals_.upload(a.values, s);
        a_vals_ptr_ = a_vals_.ptr;
    }
    dinv_.upload(inverse_diagonal, s);
    p_cols_.release();
    p_vals_.reset();
    p_starts_.release();
    ap_starts_.release();
    ap_cols_.release();
    ap_vals_.release();
    ap_host_starts_.clear();
}

void DeviceSetup::check_overflow(const char *phase) {
    u32 flag = 0;
    overflow_.download(&flag, 1, ctx_->stream);
    SF_HIP(hipStreamSynchronize(ctx_->stream));
    if (flag) throw Overflow{};
    (void)phase;
}

void DeviceSetup::clear_overflow() {
    if (!overflow_.ptr) overflow_.resize(1);
    const u32 zero = 0;
    overflow_.upload(&zero, 1, ctx_->stream);
}

double DeviceSetup::spectral(std::vector<double> &power, int iterations) {
    hipStream_t s = ctx_->stream;
    const size_t n = rows_ * (size_t)rb_;
    x_.upload(power, s);
    y_.resize(n);
    host_y_.resize(n);
    double norm = 0.0;
    for (int it = 0; it < iterations; ++it) {
        if (rb_ == 3)
            power_kernel<3><<<grid_of(rows_, 256), 256, 0, s>>>(rows_, a_starts_.ptr, a_cols_.ptr, a_vals_ptr_, dinv_.ptr, x_.ptr, y_.ptr);
        else
            power_kernel<6><<<grid_of(rows_, 256), 256, 0, s>>>(rows_, a_starts_.ptr, a_cols_.ptr, a_vals_ptr_, dinv_.ptr, x_.ptr, y_.ptr);
        SF_HIP(hipGetLastError());
        y_.download(host_y_.data(), n, s);
        SF_HIP(hipStreamSynchronize(s));
        double sum = 0.0;
        for (double v : host_y_) sum += v * v;
        norm = std::sqrt(sum);
        divide_kernel<<<grid_of(n, 256), 256, 0, s>>>(n, y_.ptr, norm, x_.ptr);
        SF_HIP(hipGetLastError());
    }
    SF_HIP(hipStreamSynchronize(s));
    return norm;
}

bool DeviceSetup::smooth(const std::vector<u32> &aggregate, u32 count, const std::vector<int64_t> &ns, const std::vector<u32> &nc,
                         const std::vector<double> &tentative, double omega, HostBsr<double> &p) {
    hipStream_t s = ctx_->stream;
    DeviceBuffer<u32> agg, ncols;
    DeviceBuffer<int64_t> nstarts;
    DeviceBuffer<double> tent;
    agg.upload(aggregate, s);
    nstarts.upload(ns, s);
    if (!nc.empty()) ncols.upload(nc, s); else ncols.resize(1);
    tent.upload(tentative, s);
    DeviceBuffer<u32> counts(rows_);
    clear_overflow();
    const SetupRows a{a_starts_.ptr, a_cols_.ptr, a_vals_ptr_};
    auto launch = [&](auto count_pass, const int64_t *out_starts, u32 *out_cols, double *out_vals) {
        constexpr bool C = decltype(count_pass)::value;
        if (rb_ == 3)
            smooth_kernel<3, C><<<rows_, 64, 0, s>>>(rows_, a, agg.ptr, nstarts.ptr, ncols.ptr, tent.ptr, dinv_.ptr, omega, counts.ptr,
                                                     out_starts, out_cols, out_vals, overflow_.ptr, slot_limit_);
        else
            smooth_kernel<6, C><<<rows_, 64, 0, s>>>(rows_, a, agg.ptr, nstarts.ptr, ncols.ptr, tent.ptr, dinv_.ptr, omega, counts.ptr,
                                                     out_starts, out_cols, out_vals, overflow_.ptr, slot_limit_);
        SF_HIP(hipGetLastError());
    };
    launch(std::true_type{}, nullptr, nullptr, nullptr);
    try {
        check_overflow("smooth");
    } catch (Overflow &) {
        return false;
    }
    std::vector<u32> host_counts(rows_);
    counts.download(host_counts.data(), rows_, s);
    SF_HIP(hipStreamSynchronize(s));
    p.rows = rows_;
    p.cols = count;
    p.rb = (size_t)rb_;
    p.cb = MODES6;
    p.starts = starts_of(host_counts);
    const size_t blocks = (size_t)p.starts.back();
    p_starts_.upload(p.starts, s);
    p_cols_.resize(std::max<size_t>(blocks, 1));
    p_vals_ = std::make_shared<DeviceBuffer<double>>(std::max<size_t>(blocks * rb_ * MODES6, 1));
    p_blocks_ = blocks;
    launch(std::false_type{}, p_starts_.ptr, p_cols_.ptr, p_vals_->ptr);
    // the pattern comes to the host; the values stay here (prolongator_values())
    p.columns.resize(blocks);
    p.values.clear();
    p_cols_.download(p.columns.data(), blocks, s);
    SF_HIP(hipStreamSynchronize(s));
    return true;
}

std::shared_ptr<const ValueSource> DeviceSetup::coarse_values() {
    if (!coarse_vals_) throw Error(SF_ERR_INTERNAL, "no coarse matrix on the device");
    return std::move(coarse_vals_);
}

std::shared_ptr<const ValueSource> DeviceSetup::prolongator_values() {
    if (!p_vals_) throw Error(SF_ERR_INTERNAL, "no prolongator on the device");
    return std::make_shared<DeviceValues>(ctx_, p_vals_, p_blocks_, (size_t)rb_ * MODES6);
}

bool DeviceSetup::product(u32 count) {
    hipStream_t s = ctx_->stream;
    DeviceBuffer<u32> counts(rows_);
    clear_overflow();
    const SetupRows a{a_starts_.ptr, a_cols_.ptr, a_vals_ptr_};
    const SetupRows p{p_starts_.ptr, p_cols_.ptr, p_vals_->ptr};
    auto launch = [&](auto count_pass, const int64_t *out_starts, u32 *out_cols, double *out_vals) {
        constexpr bool C = decltype(count_pass)::value;
        if (rb_ == 3)
            product_kernel<3, C><<<rows_, 64, 0, s>>>(rows_, a, p, counts.ptr, out_starts, out_cols, out_vals, overflow_.ptr, slot_limit_);
        else
            product_kernel<6, C><<<rows_, 64, 0, s>>>(rows_, a, p, counts.ptr, out_starts, out_cols, out_vals, overflow_.ptr, slot_limit_);
        SF_HIP(hipGetLastError());
    };
    launch(std::true_type{}, nullptr, nullptr, nullptr);
    try {
        check_overflow("product");
    } catch (Overflow &) {
        return false;
    }
    std::vector<u32> host_counts(rows_);
    counts.download(host_counts.data(), rows_, s);
    SF_HIP(hipStreamSynchronize(s));
    ap_host_starts_ = starts_of(host_counts);
    const size_t blocks = (size_t)ap_host_starts_.back();
    ap_starts_.upload(ap_host_starts_, s);
    ap_cols_.resize(std::max<size_t>(blocks, 1));
    ap_vals_.resize(std::max<size_t>(blocks * rb_ * MODES6, 1));
    launch(std::false_type{}, ap_starts_.ptr, ap_cols_.ptr, ap_vals_.ptr);
    SF_HIP(hipStreamSynchronize(s));
    (void)count;
    return true;
}

void DeviceSetup::product_values(HostBsr<double> &ap) {
    hipStream_t s = ctx_->stream;
    if (ap_host_starts_.size() != rows_ + 1) throw Error(SF_ERR_INTERNAL, "no A P on the device");
    const size_t blocks = (size_t)ap_host_starts_.back();
    ap.starts = ap_host_starts_;
    ap.columns.resize(blocks);
    ap.values.resize(blocks * rb_ * MODES6);
    ap_cols_.download(ap.columns.data(), blocks, s);
    ap_vals_.download(ap.values.data(), ap.values.size(), s);
    SF_HIP(hipStreamSynchronize(s));
}

bool DeviceSetup::galerkin(u32 count, const std::vector<int64_t> &pcol_starts, const std::vector<std::pair<u32, int64_t>> &entries,
                           HostBsr<double> &coarse) {
    hipStream_t s = ctx_->stream;
    std::vector<u32> rows(entries.size());
    std::vector<int64_t> at(entries.size());
    for (size_t i = 0; i < entries.size(); ++i) {
        rows[i] = entries[i].first;
        at[i] = entries[i].second;
    }
    DeviceBuffer<int64_t> cstarts, cat;
    DeviceBuffer<u32> crows;
    cstarts.upload(pcol_starts, s);
    if (rows.empty()) {
        rows.push_back(0);
        at.push_back(0);
    }
    crows.upload(rows, s);
    cat.upload(at, s);
    DeviceBuffer<u32> counts(std::max<size_t>(count, 1));
    clear_overflow();
    const SetupRows ap{ap_starts_.ptr, ap_cols_.ptr, ap_vals_.ptr};
    auto launch = [&](auto count_pass, const int64_t *out_starts, u32 *out_cols, double *out_vals) {
        constexpr bool C = decltype(count_pass)::value;
        if (count == 0) return;
        if (rb_ == 3)
            galerkin_kernel<3, C><<<count, 64, 0, s>>>(count, cstarts.ptr, crows.ptr, cat.ptr, p_vals_->ptr, ap, counts.ptr, out_starts,
                                                       out_cols, out_vals, overflow_.ptr, slot_limit_);
        else
            galerkin_kernel<6, C><<<count, 64, 0, s>>>(count, cstarts.ptr, crows.ptr, cat.ptr, p_vals_->ptr, ap, counts.ptr, out_starts,
                                                       out_cols, out_vals, overflow_.ptr, slot_limit_);
        SF_HIP(hipGetLastError());
    };
    launch(std::true_type{}, nullptr, nullptr, nullptr);
    try {
        check_overflow("galerkin");
    } catch (Overflow &) {
        return false;
    }
    std::vector<u32> host_counts(count);
    counts.download(host_counts.data(), count, s);
    SF_HIP(hipStreamSynchronize(s));
    coarse.rows = coarse.cols = count;
    coarse.rb = coarse.cb = MODES6;
    coarse.starts = starts_of(host_counts);
    const size_t blocks = (size_t)coarse.starts.back();
    DeviceBuffer<int64_t> out_starts;
    out_starts.upload(coarse.starts, s);
    DeviceBuffer<u32> out_cols(std::max<size_t>(blocks, 1));
    auto out_vals = std::make_shared<DeviceBuffer<double>>(std::max<size_t>(blocks * MODES6 * MODES6, 1));
    launch(std::false_type{}, out_starts.ptr, out_cols.ptr, out_vals->ptr);
    // the pattern comes to the host; the values stay here (coarse_values())
    coarse.columns.resize(blocks);
    coarse.values.clear();
    out_cols.download(coarse.columns.data(), blocks, s);
    SF_HIP(hipStreamSynchronize(s));
    coarse_vals_ = std::make_shared<DeviceValues>(ctx_, out_vals, blocks, (size_t)(MODES6 * MODES6));
    // the level is done: its device copies go
    ap_starts_.release();
    ap_cols_.release();
    ap_vals_.release();
    return true;
}

bool DeviceSetup::seed(int level, const std::vector<int64_t> &starts, const std::vector<u32> &columns, std::vector<u32> &aggregate,
                       std::vector<char> &seeded, u32 &count) {
    (void)level;
    if (!aggregate_on_device_) return false;
    const size_t rows = starts.size() - 1;
    if (rows == 0 || rows >= (size_t)UINT_MAX || columns.size() >= (size_t)UINT_MAX) return false;
    // its own stream: the power iteration may be running on the context's stream from another thread
    SF_HIP(hipSetDevice(ctx_->device));
    hipStream_t s = nullptr;
    SF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamSynchronize(s), (void)hipStreamDestroy(s); }
    } guard{s};
    std::vector<u32> st(rows + 1);
    std::vector<uint8_t> state(rows);
    for (size_t r = 0; r <= rows; ++r) st[r] = (u32)starts[r];
    for (size_t r = 0; r < rows; ++r) state[r] = starts[r + 1] > starts[r] ? SEED_UNDECIDED : SEED_NEVER;
    DeviceBuffer<u32> d_starts, d_cols, d_bid(rows), d_owner(rows), d_flag(1);
    DeviceBuffer<uint8_t> d_state;
    d_starts.upload(st, s);
    d_cols.upload(columns.empty() ? std::vector<u32>{0} : columns, s);
    d_state.upload(state, s);
    const unsigned grid = (unsigned)((rows + 255) / 256);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid), dim3(256), 0, s, rows, (u32)UINT_MAX, d_bid.ptr);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid), dim3(256), 0, s, rows, (u32)UINT_MAX, d_owner.ptr);
    constexpr int ROUNDS_PER_CHECK = 8;
    for (size_t round = 0;; round += ROUNDS_PER_CHECK) {
        if (round > rows + ROUNDS_PER_CHECK) throw Error(SF_ERR_INTERNAL, "the seed rounds did not end");
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, (size_t)1, 0u, d_flag.ptr);
        for (int i = 0; i < ROUNDS_PER_CHECK; ++i) {
            hipLaunchKernelGGL(seed_claim_kernel, dim3(grid), dim3(256), 0, s, (u32)rows, (const u32 *)d_starts.ptr,
                               (const u32 *)d_cols.ptr, (const uint8_t *)d_state.ptr, d_bid.ptr);
            hipLaunchKernelGGL(seed_decide_kernel, dim3(grid), dim3(256), 0, s, (u32)rows, (const u32 *)d_starts.ptr,
                               (const u32 *)d_cols.ptr, d_state.ptr, (const u32 *)d_bid.ptr, d_owner.ptr);
            hipLaunchKernelGGL(seed_settle_kernel, dim3(grid), dim3(256), 0, s, (u32)rows, (const u32 *)d_starts.ptr,
                               (const u32 *)d_cols.ptr, d_state.ptr, (const u32 *)d_owner.ptr, d_bid.ptr,
                               i == ROUNDS_PER_CHECK - 1 ? d_flag.ptr : nullptr);
        }
        SF_HIP(hipGetLastError());
        u32 left = 0;
        d_flag.download(&left, 1, s);
        SF_HIP(hipStreamSynchronize(s));
        if (!left) break;
    }
    std::vector<u32> owner(rows);
    d_owner.download(owner.data(), rows, s);
    d_state.download(state.data(), rows, s);
    SF_HIP(hipStreamSynchronize(s));
    // aggregate numbers in seed order, which is index order
    std::vector<u32> id(rows, UINT_MAX);
    for (size_t r = 0; r < rows; ++r)
        if (state[r] == SEED_IN) id[r] = count++;
    for (size_t r = 0; r < rows; ++r)
        if (owner[r] != UINT_MAX) {
            aggregate[r] = id[owner[r]];
            seeded[r] = 1;
        }
    return true;
}

bool DeviceSetup::factor(int level, size_t n, double *a) {
    // below this size the host's column loop is as quick as the launches
    if (n < 128 || n > 65535) return false;
    SF_HIP(hipSetDevice(ctx_->device));
    hipStream_t s = nullptr;  // its own: the power iteration may be running on the context's stream
    SF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamSynchronize(s), (void)hipStreamDestroy(s); }
    } guard{s};
    DeviceBuffer<double> d_a, d_pivots(n);
    d_a.upload(a, n * n, s);
    for (size_t k = 0; k < n; ++k) {
        const size_t rest = n - k;  // rows k .. n - 1
        hipLaunchKernelGGL(cholesky_finish_column_kernel, dim3((unsigned)((rest + 255) / 256)), dim3(256), 0, s, (u32)n, (u32)k, d_a.ptr,
                           d_pivots.ptr);
        if (rest > 1)
            hipLaunchKernelGGL(cholesky_update_kernel, dim3((unsigned)((rest - 1 + 255) / 256), (unsigned)(rest - 1)), dim3(256), 0, s,
                               (u32)n, (u32)k, d_a.ptr);
    }
    hipLaunchKernelGGL(cholesky_diagonal_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (u32)n, (const double *)d_pivots.ptr,
                       d_a.ptr);
    SF_HIP(hipGetLastError());
    std::vector<double> pivots(n);
    d_pivots.download(pivots.data(), n, s);
    d_a.download(a, n * n, s);
    SF_HIP(hipStreamSynchronize(s));
    for (size_t j = 0; j < n; ++j)
        if (!std::isfinite(pivots[j]) || pivots[j] <= 0.0)
            throw Error(SF_ERR_MG_NOT_POSITIVE_DEFINITE, "the coarse matrix is not positive definite", level, (int64_t)j, pivots[j]);
    return true;
}

void DeviceSetup::end_level() {
    a_starts_.release();
    a_cols_.release();
    a_vals_.release();
    a_vals_ptr_ = nullptr;
    dinv_.release();
    p_starts_.release();
    p_cols_.release();
    p_vals_.reset();
    ap_starts_.release();
    ap_cols_.release();
    ap_vals_.release();
    ap_host_starts_.clear();
    x_.release();
    y_.release();
    host_y_.clear();
    host_y_.shrink_to_fit();
}

DeviceValues::DeviceValues(std::shared_ptr<Context> ctx, const double *values, size_t blocks, size_t block_len)
    : ctx_(std::move(ctx)), values_(values), blocks_(blocks), len_(block_len) {}

DeviceValues::DeviceValues(std::shared_ptr<Context> ctx, std::shared_ptr<DeviceBuffer<double>> owned, size_t blocks, size_t block_len)
    : ctx_(std::move(ctx)), owned_(std::move(owned)), values_(owned_->ptr), blocks_(blocks), len_(block_len) {}

void DeviceValues::values(double *out) const {
    hipStream_t s = ctx_->stream;
    if (blocks_) SF_HIP(hipMemcpyAsync(out, values_, blocks_ * len_ * sizeof(double), hipMemcpyDeviceToHost, s));
    SF_HIP(hipStreamSynchronize(s));
}

void DeviceValues::blocks(const std::vector<int64_t> &slots, double *out) const {
    if (slots.empty()) return;
    for (int64_t at : slots)
        if (at < 0 || (size_t)at >= blocks_) throw Error(SF_ERR_INTERNAL, "a block slot outside the matrix");
    hipStream_t s = ctx_->stream;
    DeviceBuffer<int64_t> d_slots;
    d_slots.upload(slots, s);
    const size_t n = slots.size() * len_;
    DeviceBuffer<double> d_out(n);
    gather_blocks_kernel<<<grid_of(n, 256), 256, 0, s>>>(slots.size(), (int)len_, d_slots.ptr, values_, d_out.ptr);
    SF_HIP(hipGetLastError());
    d_out.download(out, n, s);
    SF_HIP(hipStreamSynchronize(s));
}

void DeviceValues::block_squares(double *out) const {
    if (!blocks_) return;
    hipStream_t s = ctx_->stream;
    DeviceBuffer<double> d_out(blocks_);
    block_squares_kernel<<<grid_of(blocks_, 256), 256, 0, s>>>(blocks_, (int)len_, values_, d_out.ptr);
    SF_HIP(hipGetLastError());
    d_out.download(out, blocks_, s);
    SF_HIP(hipStreamSynchronize(s));
}

}  // namespace sf
