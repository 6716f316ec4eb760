// Device objects of the solve path: sliced matrices, the V-cycle, the conjugate gradient.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>

#include "sf_batch_kernels.hpp"
#include "sf_device.hpp"
#include "sf_setup_device.hpp"

namespace sf {

using namespace dev;

static inline unsigned grid_for(size_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

Context::Context(int dev) : device(dev) {
    int count = 0;
    SF_HIP(hipGetDeviceCount(&count));
    if (dev < 0 || dev >= count) throw Error(SF_ERR_INVALID_ARGUMENT, "no such device");
    SF_HIP(hipSetDevice(dev));
    SF_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
}
Context::~Context() {
    if (stream) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
    }
}

template <typename T>
uint32_t DeviceSliced<T>::xcd_order() {
    static const uint32_t order = [] {
        // whole sort windows per XCD by default: measured 11% faster on the plate's Chebyshev step
        const char *e = std::getenv("SOLIDFEM_XCD_ORDER");
        const int v = e ? std::atoi(e) : 2;
        return v == 1 ? (uint32_t)ORDER_XCD_RUNS : v == 2 ? (uint32_t)ORDER_XCD_WINDOWS : 0u;
    }();
    return order;
}
template <typename T>
int DeviceSliced<T>::group_runs() {
    // SOLIDFEM_GROUP_RUNS: 0 (default) dispatch order, 1 one contiguous run per XCD, a power of two
    // W > 1 chunks of W workgroups per XCD in turn
    static const int on = [] {
        const char *e = std::getenv("SOLIDFEM_GROUP_RUNS");
        const int v = e ? std::atoi(e) : 0;
        return (v > 1 && (v & (v - 1)) == 0 && v <= 1024) || v == 1 ? v : 0;
    }();
    return on;
}
template <typename T>
bool DeviceBsr<T>::rows_ping_pong() {
    // on unless SOLIDFEM_ROWS_PING_PONG=0: measured 0.6-1.1 ms faster per plate FP32 solve
    static const bool on = [] {
        const char *e = std::getenv("SOLIDFEM_ROWS_PING_PONG");
        return !e || std::atoi(e) != 0;
    }();
    return on;
}
template struct DeviceBsr<double>;
template struct DeviceBsr<float>;

template <typename T>
bool DeviceSliced<T>::ping_pong() {
    static const bool on = [] {
        const char *e = std::getenv("SOLIDFEM_PING_PONG");
        return e && std::atoi(e) != 0;
    }();
    return on;
}
template struct DeviceSliced<double>;
template struct DeviceSliced<float>;

// ---- launch timing -----------------------------------------------------------------------------

Profiler &Profiler::get() {
    static Profiler p;
    return p;
}

int Profiler::begin(int kernel, hipStream_t s, double bytes, u64 launches) {
    if (kernel < 0 || kernel >= KERNELS) return -1;
    std::lock_guard<std::mutex> g(lock);
    Slot &slot = slots[kernel];
    if (!slot.on) return -1;
    Pending p{};
    for (hipEvent_t *e : {&p.start, &p.stop}) {
        if (!slot.pool.empty()) {
            *e = slot.pool.back();
            slot.pool.pop_back();
        } else {
            SF_HIP(hipEventCreate(e));
        }
    }
    p.bytes = bytes;
    p.launches = launches;
    SF_HIP(hipEventRecord(p.start, s));
    slot.pending.push_back(p);
    return (int)slot.pending.size() - 1;
}

void Profiler::end(int kernel, int token, hipStream_t s) {
    if (token < 0) return;
    std::lock_guard<std::mutex> g(lock);
    SF_HIP(hipEventRecord(slots[kernel].pending[(size_t)token].stop, s));
}

void Profiler::collect(int kernel) {
    std::lock_guard<std::mutex> g(lock);
    Slot &slot = slots[kernel];
    for (auto &p : slot.pending) {
        SF_HIP(hipEventSynchronize(p.stop));
        float ms = 0;
        SF_HIP(hipEventElapsedTime(&ms, p.start, p.stop));
        slot.seconds += ms * 1e-3;
        slot.bytes += p.bytes;
        slot.launches += p.launches;
        slot.pool.push_back(p.start);
        slot.pool.push_back(p.stop);
    }
    slot.pending.clear();
}

void Profiler::reset() {
    for (int k = 0; k < KERNELS; ++k) {
        collect(k);
        std::lock_guard<std::mutex> g(lock);
        slots[k].launches = 0;
        slots[k].seconds = slots[k].bytes = 0;
    }
}

// ---- sliced layout ---------------------------------------------------------------------------

SlicedPlan plan_sliced(const int64_t *starts, const u32 *columns, size_t rows, int rb, int cb) {
    SlicedPlan p;
    p.rb = rb;
    p.cb = cb;
    p.npw = WAVE / rb;
    p.rows = rows;
    p.blocks = (size_t)starts[rows];
    const size_t npw = (size_t)p.npw;
    p.slices = (uint32_t)((rows + npw - 1) / npw);
    // Rows sorted by block count (descending, stable) inside windows of 32 slices: a slice's
    // lanes then run loops of nearly equal length while rows stay near their neighbours. Windows
    // and slices are independent, so they are laid out on the host pool.
    std::vector<uint32_t> order(rows);
    const size_t window = 32 * npw;
    const size_t windows = (rows + window - 1) / window;
    parallel_for(windows, 8, [&](size_t first, size_t last) {
        for (size_t w = first; w < last; ++w) {
            const size_t w0 = w * window, w1 = std::min(rows, w0 + window);
            std::iota(order.begin() + w0, order.begin() + w1, (uint32_t)w0);
            std::stable_sort(order.begin() + w0, order.begin() + w1, [&](uint32_t a, uint32_t b) {
                return (starts[a + 1] - starts[a]) > (starts[b + 1] - starts[b]);
            });
        }
    });
    p.perm.assign((size_t)p.slices * npw, NONE);
    p.row_len.assign((size_t)p.slices * npw, 0);
    p.slice_start.assign((size_t)p.slices + 1, 0);
    parallel_for(p.slices, 1024, [&](size_t first, size_t last) {
        for (size_t s = first; s < last; ++s) {
            uint32_t width = 0;
            for (size_t j = 0; j < npw; ++j) {
                const size_t at = s * npw + j;
                if (at >= rows) break;
                const uint32_t row = order[at];
                p.perm[at] = row;
                p.row_len[at] = (uint32_t)(starts[row + 1] - starts[row]);
                width = std::max(width, p.row_len[at]);
            }
            p.slice_start[s + 1] = width;
        }
    });
    for (uint32_t s = 0; s < p.slices; ++s) p.slice_start[s + 1] += p.slice_start[s];
    const size_t slots = p.slice_start[p.slices];
    p.cols.assign(slots * npw, 0);
    p.block_of_slot.assign(slots * npw, NONE);
    parallel_for(p.slices, 1024, [&](size_t first, size_t last) {
        for (size_t s = first; s < last; ++s)
            for (size_t j = 0; j < npw; ++j) {
                const size_t at = s * npw + j;
                if (p.perm[at] == NONE) continue;
                const uint32_t row = p.perm[at];
                for (uint32_t k = 0; k < p.row_len[at]; ++k) {
                    const size_t slot = (size_t)(p.slice_start[s] + k) * npw + j;
                    p.cols[slot] = columns[(size_t)starts[row] + k];
                    p.block_of_slot[slot] = (uint32_t)(starts[row] + k);
                }
            }
    });
    return p;
}

template <typename T>
void build_sliced(DeviceSliced<T> &out, const SlicedPlan &plan, size_t block_cols, const double *src, hipStream_t s) {
    out.rb = plan.rb;
    out.cb = plan.cb;
    out.rows = plan.rows;
    out.block_cols = block_cols;
    out.blocks = plan.blocks;
    out.slices = plan.slices;
    out.slice_start.upload(plan.slice_start, s);
    out.row_len.upload(plan.row_len, s);
    out.perm.upload(plan.perm, s);
    out.cols.upload(plan.cols, s);
    DeviceBuffer<uint32_t> map;
    map.upload(plan.block_of_slot, s);
    const size_t n = (size_t)plan.slice_start[plan.slices] * plan.cb * WAVE;
    out.vals.resize(std::max<size_t>(n, 1));
    if (n) hipLaunchKernelGGL(gather_convert_kernel<T>, dim3(grid_for(n, 256)), dim3(256), 0, s, n, map.ptr, plan.rb, plan.cb, src, out.vals.ptr);
    SF_HIP(hipGetLastError());
    SF_HIP(hipStreamSynchronize(s));  // `map` is released on return
}

template <typename T>
void upload_sliced(DeviceSliced<T> &out, const HostBsr<double> &m, hipStream_t s) {
    SlicedPlan plan = plan_sliced(m.starts.data(), m.columns.data(), m.rows, (int)m.rb, (int)m.cb);
    DeviceBuffer<double> src;
    src.upload(m.values, s);
    build_sliced(out, plan, m.cols, src.ptr, s);
}

// The same sliced matrix in precision T from one already on the device in f64: the layout arrays
// are copied and every value rounded to T, which is what the host path's gather would produce.
template <typename T>
void convert_sliced(DeviceSliced<T> &out, const DeviceSliced<double> &in, hipStream_t s) {
    out.rb = in.rb;
    out.cb = in.cb;
    out.rows = in.rows;
    out.block_cols = in.block_cols;
    out.blocks = in.blocks;
    out.slices = in.slices;
    auto copy = [&](DeviceBuffer<uint32_t> &to, const DeviceBuffer<uint32_t> &from) {
        to.resize(from.count);
        if (from.count) SF_HIP(hipMemcpyAsync(to.ptr, from.ptr, from.count * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    };
    copy(out.slice_start, in.slice_start);
    copy(out.row_len, in.row_len);
    copy(out.perm, in.perm);
    copy(out.cols, in.cols);
    out.vals.resize(in.vals.count);
    if (in.vals.count)
        hipLaunchKernelGGL((convert_kernel<T, double>), dim3(grid_for(in.vals.count, 256)), dim3(256), 0, s, in.vals.count,
                           (const double *)in.vals.ptr, out.vals.ptr);
    SF_HIP(hipGetLastError());
}

template void upload_sliced<double>(DeviceSliced<double> &, const HostBsr<double> &, hipStream_t);
template void upload_sliced<float>(DeviceSliced<float> &, const HostBsr<double> &, hipStream_t);
template void build_sliced<double>(DeviceSliced<double> &, const SlicedPlan &, size_t, const double *, hipStream_t);
template void build_sliced<float>(DeviceSliced<float> &, const SlicedPlan &, size_t, const double *, hipStream_t);

template <int RB, int CB, int BATCH, typename S, typename V>
static void spmv_batch(const DeviceSliced<S> &a, const V *x, V *y, const V *b, int epi, hipStream_t s, Start<S, V> st) {
    const dim3 grid(grid_for(a.slices, 4)), block(256);
    if (epi == EPI_STORE) hipLaunchKernelGGL((spmv_kernel<RB, CB, BATCH, S, V, EPI_STORE>), grid, block, 0, s, a.view(), x, y, b, st);
    else if (epi == EPI_RESIDUAL) hipLaunchKernelGGL((spmv_kernel<RB, CB, BATCH, S, V, EPI_RESIDUAL>), grid, block, 0, s, a.view(), x, y, b, st);
    else hipLaunchKernelGGL((spmv_kernel<RB, CB, BATCH, S, V, EPI_ADD>), grid, block, 0, s, a.view(), x, y, b, st);
}

template <int RB, int CB, typename S, typename V>
static void spmv_shape(const DeviceSliced<S> &a, const V *x, V *y, const V *b, int epi, hipStream_t s, int batch, Start<S, V> st) {
    switch (batch) {
        case 1: spmv_batch<RB, CB, 1>(a, x, y, b, epi, s, st); break;
        case 2: spmv_batch<RB, CB, 2>(a, x, y, b, epi, s, st); break;
        case 8: spmv_batch<RB, CB, 8>(a, x, y, b, epi, s, st); break;
        default: spmv_batch<RB, CB, 4>(a, x, y, b, epi, s, st); break;
    }
}

template <typename S, typename V>
void launch_spmv(const DeviceSliced<S> &a, const V *x, V *y, const V *b, int epi, hipStream_t s, int batch, Start<S, V> st) {
    if (a.slices == 0) return;
    if (batch == 0) batch = a.batch_spmv;
    if (st.dinv && a.rb != 3 && a.rb != 6) throw Error(SF_ERR_INTERNAL, "a fused start needs 3 or 6 rows per block");
    if (a.rb == 3 && a.cb == 3) spmv_shape<3, 3>(a, x, y, b, epi, s, batch, st);
    else if (a.rb == 6 && a.cb == 6) spmv_shape<6, 6>(a, x, y, b, epi, s, batch, st);
    else if (a.rb == 3 && a.cb == 6) spmv_shape<3, 6>(a, x, y, b, epi, s, batch, st);
    else if (a.rb == 6 && a.cb == 3) spmv_shape<6, 3>(a, x, y, b, epi, s, batch, st);
    else if constexpr (std::is_same<S, double>::value && std::is_same<V, double>::value) {
        if (a.rb == 1 && a.cb == 1) spmv_shape<1, 1>(a, x, y, b, epi, s, batch, st);
        else throw Error(SF_ERR_INTERNAL, "unsupported block shape");
    } else {
        throw Error(SF_ERR_INTERNAL, "unsupported block shape");
    }
    SF_HIP(hipGetLastError());
}

template <typename T>
void upload_bsr(DeviceBsr<T> &out, const HostBsr<double> &m, hipStream_t s) {
    if (m.starts.back() > (int64_t)UINT32_MAX) throw Error(SF_ERR_INTERNAL, "a block matrix too large for 32-bit offsets");
    out.rb = (int)m.rb;
    out.cb = (int)m.cb;
    out.rows = (uint32_t)m.rows;
    std::vector<uint32_t> starts(m.starts.begin(), m.starts.end());
    std::vector<T> vals(m.values.size());
    parallel_for(vals.size(), 65536, [&](size_t first, size_t last) {
        for (size_t i = first; i < last; ++i) vals[i] = (T)m.values[i];
    });
    out.starts.upload(starts, s);
    out.cols.upload(m.columns.empty() ? std::vector<uint32_t>{0} : m.columns, s);
    out.vals.upload(vals.empty() ? std::vector<T>{T(0)} : vals, s);
}
// Block k of `out` (cb x rb, row-major) is the transpose of block from[k] of `in` (rb x cb).
static __global__ void transpose_blocks_kernel(size_t blocks, int rb, int cb, const uint32_t *__restrict__ from,
                                        const double *__restrict__ in, double *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t len = (size_t)rb * (size_t)cb;
    if (i >= blocks * len) return;
    const size_t k = i / len, e = i % len;
    const size_t q = e / (size_t)rb, p = e % (size_t)rb;
    out[i] = in[(size_t)from[k] * len + p * (size_t)cb + q];
}

template <typename T>
void build_bsr(DeviceBsr<T> &out, const HostBsr<double> &pattern, const double *values, hipStream_t s) {
    if (pattern.starts.back() > (int64_t)UINT32_MAX) throw Error(SF_ERR_INTERNAL, "a block matrix too large for 32-bit offsets");
    out.rb = (int)pattern.rb;
    out.cb = (int)pattern.cb;
    out.rows = (uint32_t)pattern.rows;
    std::vector<uint32_t> starts(pattern.starts.begin(), pattern.starts.end());
    out.starts.upload(starts, s);
    out.cols.upload(pattern.columns.empty() ? std::vector<uint32_t>{0} : pattern.columns, s);
    const size_t n = pattern.blocks() * pattern.block_len();
    out.vals.resize(std::max<size_t>(n, 1));
    if (n) hipLaunchKernelGGL((convert_kernel<T, double>), dim3(grid_for(n, 256)), dim3(256), 0, s, n, values, out.vals.ptr);
    SF_HIP(hipGetLastError());
    SF_HIP(hipStreamSynchronize(s));
}
template void build_bsr<double>(DeviceBsr<double> &, const HostBsr<double> &, const double *, hipStream_t);
template void build_bsr<float>(DeviceBsr<float> &, const HostBsr<double> &, const double *, hipStream_t);
template void upload_bsr<double>(DeviceBsr<double> &, const HostBsr<double> &, hipStream_t);
template void upload_bsr<float>(DeviceBsr<float> &, const HostBsr<double> &, hipStream_t);

template <int RB, int CB, typename S, typename V, int RPW>
static void long_rows_epi(const DeviceBsr<S> &a, const V *x, V *y, const V *b, int epi, hipStream_t s, Start<S, V> st) {
    const dim3 grid((a.rows + RPW - 1) / RPW), block(WAVE);
    if (epi == EPI_STORE) hipLaunchKernelGGL((long_rows_kernel<RB, CB, S, V, EPI_STORE, RPW>), grid, block, 0, s, a.view(), x, y, b, st);
    else if (epi == EPI_RESIDUAL)
        hipLaunchKernelGGL((long_rows_kernel<RB, CB, S, V, EPI_RESIDUAL, RPW>), grid, block, 0, s, a.view(), x, y, b, st);
    else hipLaunchKernelGGL((long_rows_kernel<RB, CB, S, V, EPI_ADD, RPW>), grid, block, 0, s, a.view(), x, y, b, st);
}

template <int RB, int CB, typename S, typename V>
static void long_rows_shape(const DeviceBsr<S> &a, const V *x, V *y, const V *b, int epi, hipStream_t s, Start<S, V> st, int rpw) {
    if (rpw == 4) long_rows_epi<RB, CB, S, V, 4>(a, x, y, b, epi, s, st);
    else if (rpw == 2) long_rows_epi<RB, CB, S, V, 2>(a, x, y, b, epi, s, st);
    else long_rows_epi<RB, CB, S, V, 1>(a, x, y, b, epi, s, st);
}

template <typename S, typename V>
void launch_long_rows(const DeviceBsr<S> &a, const V *x, V *y, const V *b, int epi, hipStream_t s, Start<S, V> st, int rpw) {
    if (a.rows == 0) return;
    if (a.rb == 6 && a.cb == 3) long_rows_shape<6, 3>(a, x, y, b, epi, s, st, rpw);
    else if (a.rb == 6 && a.cb == 6) long_rows_shape<6, 6>(a, x, y, b, epi, s, st, rpw);
    else if (a.rb == 3 && a.cb == 3) long_rows_shape<3, 3>(a, x, y, b, epi, s, st, rpw);
    else if (a.rb == 3 && a.cb == 6) long_rows_shape<3, 6>(a, x, y, b, epi, s, st, rpw);
    else throw Error(SF_ERR_INTERNAL, "unsupported block shape");
    SF_HIP(hipGetLastError());
}
template void launch_long_rows<double, double>(const DeviceBsr<double> &, const double *, double *, const double *, int, hipStream_t,
                                               Start<double, double>, int);
template void launch_long_rows<float, double>(const DeviceBsr<float> &, const double *, double *, const double *, int, hipStream_t,
                                              Start<float, double>, int);
template void launch_long_rows<float, float>(const DeviceBsr<float> &, const float *, float *, const float *, int, hipStream_t,
                                             Start<float, float>, int);

std::atomic<int> &forced_variant() {
    // SOLIDFEM_BATCH=code gives the starting value; sf_force_variant changes it. The environment
    // is read once: a setenv by another thread while this thread reads it is undefined behaviour.
    static std::atomic<int> code{[] {
        const char *e = std::getenv("SOLIDFEM_BATCH");
        return e ? std::atoi(e) : 0;
    }()};
    return code;
}

int tune_batch(hipStream_t s, const std::function<void(int)> &launch, const std::vector<int> &candidates, const std::string &what,
               const std::string &memo) {
    // a forced code is taken wherever it is a candidate
    const int forced = forced_variant().load(std::memory_order_relaxed);
    for (int c : candidates)
        if (forced > 0 && forced == c) return forced;
    return fastest(s, launch, candidates, what, memo);
}

static int fastest_timed(hipStream_t s, const std::function<void(int)> &launch, const std::vector<int> &candidates,
                         const std::string &what);

namespace {
std::mutex g_memo_lock;
std::map<std::string, int> &tuning_memo() {
    static std::map<std::string, int> memo;
    return memo;
}
}  // namespace

int fastest(hipStream_t s, const std::function<void(int)> &launch, const std::vector<int> &candidates, const std::string &what,
            const std::string &memo) {
    static const bool verbose = std::getenv("SOLIDFEM_TUNE_LOG") != nullptr;
    std::string key;
    if (!memo.empty()) {
        key = memo;
        for (int c : candidates) key += "," + std::to_string(c);
        std::lock_guard<std::mutex> g(g_memo_lock);
        auto it = tuning_memo().find(key);
        if (it != tuning_memo().end()) {
            if (verbose) std::fprintf(stderr, "[tune] %s: kept %d for %s\n", what.c_str(), it->second, memo.c_str());
            return it->second;
        }
    }
    const int best = fastest_timed(s, launch, candidates, what);
    if (!key.empty()) {
        std::lock_guard<std::mutex> g(g_memo_lock);
        tuning_memo().emplace(key, best);
    }
    return best;
}

static int fastest_timed(hipStream_t s, const std::function<void(int)> &launch, const std::vector<int> &candidates,
                         const std::string &what) {
    hipEvent_t e0, e1;
    SF_HIP(hipEventCreate(&e0));
    SF_HIP(hipEventCreate(&e1));
    static const bool verbose = std::getenv("SOLIDFEM_TUNE_LOG") != nullptr;
    // two rounds over the candidates, each candidate's best kept: a drift in the GPU's state
    // during the first round (clocks, what the caches hold) does not decide the choice
    std::vector<float> least(candidates.size(), 1e30f);
    for (int round = 0; round < 2; ++round)
        for (size_t c = 0; c < candidates.size(); ++c) {
            const int batch = candidates[c];
            launch(batch);  // warm
            SF_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < 3; ++r) launch(batch);
            SF_HIP(hipEventRecord(e1, s));
            SF_HIP(hipEventSynchronize(e1));
            float ms = 0;
            SF_HIP(hipEventElapsedTime(&ms, e0, e1));
            least[c] = std::min(least[c], ms);
        }
    int best = candidates.front();
    float best_ms = 1e30f;
    for (size_t c = 0; c < candidates.size(); ++c) {
        if (verbose) std::fprintf(stderr, "[tune] %s: variant %d: %.2f us\n", what.c_str(), candidates[c], least[c] * 1e3f / 3);
        if (least[c] < best_ms) {
            best_ms = least[c];
            best = candidates[c];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (verbose) std::fprintf(stderr, "[tune] %s: chose %d\n", what.c_str(), best);
    return best;
}

template <typename S, typename V>
void tune_spmv(DeviceSliced<S> &a, int epi, hipStream_t s) {
    if (a.slices == 0) return;
    DeviceBuffer<V> x(std::max<size_t>(a.block_cols * a.cb, 1)), y(std::max<size_t>(a.rows * a.rb, 1));
    SF_HIP(hipMemsetAsync(x.ptr, 0, x.bytes(), s));
    SF_HIP(hipMemsetAsync(y.ptr, 0, y.bytes(), s));
    a.batch_spmv = tune_batch(s, [&](int batch) { launch_spmv<S, V>(a, x.ptr, y.ptr, y.ptr, epi, s, batch); }, {1, 2, 4, 8}, "",
                              "product " + std::to_string(epi) + " " + shape_class(a) + "|" + std::to_string(sizeof(V)));
    SF_HIP(hipStreamSynchronize(s));
}

template void launch_spmv<double, double>(const DeviceSliced<double> &, const double *, double *, const double *, int, hipStream_t, int,
                                          Start<double, double>);
template void launch_spmv<float, double>(const DeviceSliced<float> &, const double *, double *, const double *, int, hipStream_t, int,
                                         Start<float, double>);
template void launch_spmv<float, float>(const DeviceSliced<float> &, const float *, float *, const float *, int, hipStream_t, int,
                                        Start<float, float>);
template void tune_spmv<double, double>(DeviceSliced<double> &, int, hipStream_t);

// ---- the device hierarchy -------------------------------------------------------------------

namespace {

// coarse_inverse_small_kernel for n <= COARSE_SMALL rows. Its LDS limit is raised once per
// process, when a hierarchy forms its inverse (so never under a stream capture).
template <typename V>
void raise_coarse_small() {
    static const bool raised = [] {
        const int most = (int)coarse_small_lds(COARSE_SMALL);
        SF_HIP(hipFuncSetAttribute((const void *)coarse_inverse_small_kernel<V>, hipFuncAttributeMaxDynamicSharedMemorySize, most));
        return true;
    }();
    (void)raised;
}
template <typename V>
void launch_coarse_small(uint32_t n, const double *inv, const V *b, V *x, hipStream_t s) {
    hipLaunchKernelGGL(coarse_inverse_small_kernel<V>, dim3(1), dim3(COARSE_SMALL_THREADS), coarse_small_lds(n), s, n, inv, b, x);
}

// Chebyshev scalars computed in f64 as the reference's cycle does (cycle.rs:578-650,
// crates/solidfem-solve/src/gpu/layout.rs:445-468).
void chebyshev(double lmin, double lmax, size_t degree, double &theta, std::vector<std::pair<double, double>> &steps) {
    theta = 0.5 * (lmax + lmin);
    const double delta = 0.5 * (lmax - lmin);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;
    steps.clear();
    for (size_t s = 1; s < degree; ++s) {
        const double next = 1.0 / (2.0 * sigma - rho);
        steps.push_back({next * rho, 2.0 * next / delta});
        rho = next;
    }
}

template <typename S, typename V>
struct LevelDev {
    int rb = 3;
    size_t rows = 0, n = 0;
    DeviceSliced<S> a, p, r;
    DeviceBsr<S> r_rows;          // the same transpose in plain block rows, for long_rows_kernel
    int restrict_rows = 0;  // 0: the sliced restriction; else long_rows_kernel with this many rows per wave (by timing)
    DeviceBsr<S> a_rows;          // coarse levels: A in plain block rows, for the row kernels
    int cheb_rows = 0, residual_rows = 0;  // the same choice for the steps and residuals (by timing)
    DeviceBuffer<S> dinv;
    DeviceBuffer<double> lt;      // coarsest: transposed Cholesky factor
    DeviceBuffer<double> packed;  // coarsest: the factor packed by columns, for the LDS solve
    DeviceBuffer<double> inverse; // coarsest: the f64 inverse, by columns (SF_COARSE_INVERSE)
    size_t lds_bytes = 0;         // nonzero when the LDS solve is used
    bool has_p = false, has_factor = false;
    double theta = 0;
    std::vector<std::pair<double, double>> steps;
    DeviceBuffer<V> b, x, d0, d1, r0, r1, res, px;
    // the same work vectors for several columns (row-interleaved), grown on demand
    DeviceBuffer<V> mb, mx, md0, md1, mr0, mr1, mres, mpx;
    // batch depths of the batched kernels per log2 of the column group: Chebyshev step, residual,
    // restriction, prolongation (0 = not tuned yet)
    int mm_depth[8][4] = {};
};

template <typename S, typename V>
class Hierarchy final : public DeviceMultigrid {
   public:
    struct Graph {
        const double *r;
        double *z;
        hipGraphExec_t exec;
        hipEvent_t done;  // recorded after every replay of this graph
    };
    std::vector<Graph> graphs_;

    Hierarchy(std::vector<HostLevel> levels, size_t deg, int prec, int coarse, hipStream_t s, const DeviceSliced<double> *finest) {
        precision = prec;
        coarse_mode = coarse;
        degree = deg;
        host = std::move(levels);
        lv_.resize(host.size());
        StageClock clock;
        for (size_t l = 0; l < host.size(); ++l) {
            const HostLevel &h = host[l];
            LevelDev<S, V> &d = lv_[l];
            d.rb = (int)h.a.rb;
            d.rows = h.a.rows;
            d.n = h.a.dofs();
            // the matrix's values on the device where the setup left them (a coarse matrix built
            // there), else uploaded
            const double *adev = h.a.values.empty() && h.source ? h.source->device() : nullptr;
            if (l == 0 && finest && finest->rows == h.a.rows && finest->blocks == h.a.blocks() && finest->rb == (int)h.a.rb &&
                finest->cb == (int)h.a.cb)
                convert_sliced(d.a, *finest, s);  // the study's matrix, already on the device
            else if (adev)
                build_sliced(d.a, plan_sliced(h.a.starts.data(), h.a.columns.data(), h.a.rows, (int)h.a.rb, (int)h.a.cb), h.a.cols, adev, s);
            else {
                h.fetch_values();
                upload_sliced(d.a, h.a, s);
            }
            if (l > 0) {
                if (adev) build_bsr(d.a_rows, h.a, adev, s);
                else upload_bsr(d.a_rows, h.a, s);
            }
            std::vector<S> dinv(h.inverse_diagonal.size());
            for (size_t i = 0; i < dinv.size(); ++i) dinv[i] = (S)h.inverse_diagonal[i];
            d.dinv.upload(dinv, s);
            SF_HIP(hipStreamSynchronize(s));
            clock.lap("  level matrix and block inverse");
            chebyshev(h.lambda_min, h.lambda_max, degree, d.theta, d.steps);
            if (h.has_prolongator) {
                d.has_p = true;
                const auto &P = h.prolongator;
                // P's values: on the device where the setup built them, else uploaded
                const double *pdev = P.values.empty() && h.prolongator_source ? h.prolongator_source->device() : nullptr;
                DeviceBuffer<double> uploaded;
                if (!pdev) {
                    h.fetch_values();
                    uploaded.upload(P.values, s);
                    pdev = uploaded.ptr;
                }
                build_sliced(d.p, plan_sliced(P.starts.data(), P.columns.data(), P.rows, (int)P.rb, (int)P.cb), P.cols, pdev, s);
                clock.lap("  prolongator layout");
                // the stored transpose: block row c holds block column c of P in ascending fine
                // row order, each block transposed (crates/solidfem-solve/src/gpu/layout.rs:254-291);
                // its pattern on the host, its values gathered on the device
                if (P.blocks() > (size_t)UINT32_MAX) throw Error(SF_ERR_INTERNAL, "a prolongator too large for 32-bit block indices");
                HostBsr<double> t;
                t.rows = P.cols;
                t.cols = P.rows;
                t.rb = P.cb;
                t.cb = P.rb;
                t.starts.assign(t.rows + 1, 0);
                for (u32 c : P.columns) t.starts[(size_t)c + 1] += 1;
                for (size_t i = 0; i < t.rows; ++i) t.starts[i + 1] += t.starts[i];
                std::vector<int64_t> fill(t.starts.begin(), t.starts.end() - 1);
                std::vector<uint32_t> from(P.blocks());
                t.columns.assign(P.blocks(), 0);
                for (size_t row = 0; row < P.rows; ++row)
                    for (int64_t at = P.starts[row]; at < P.starts[row + 1]; ++at) {
                        const size_t to = (size_t)fill[P.columns[(size_t)at]]++;
                        t.columns[to] = (u32)row;
                        from[to] = (uint32_t)at;
                    }
                const size_t tn = P.blocks() * P.block_len();
                DeviceBuffer<double> tvals(std::max<size_t>(tn, 1));
                {
                    DeviceBuffer<uint32_t> dfrom;
                    dfrom.upload(from, s);
                    if (tn)
                        hipLaunchKernelGGL(transpose_blocks_kernel, dim3(grid_for(tn, 256)), dim3(256), 0, s, P.blocks(), (int)P.rb,
                                           (int)P.cb, dfrom.ptr, pdev, tvals.ptr);
                    SF_HIP(hipGetLastError());
                    SF_HIP(hipStreamSynchronize(s));
                }
                clock.lap("  restriction transpose");
                build_sliced(d.r, plan_sliced(t.starts.data(), t.columns.data(), t.rows, (int)t.rb, (int)t.cb), t.cols, tvals.ptr, s);
                clock.lap("  restriction layout");
                build_bsr(d.r_rows, t, tvals.ptr, s);
                clock.lap("  restriction rows");
            }
            if (h.has_factor) {
                d.has_factor = true;
                const size_t n = d.n;
                if (n > COARSE_MAX) throw Error(SF_ERR_INTERNAL, "coarse factor larger than the device solve");
                std::vector<double> lt(n * n, 0.0);
                for (size_t i = 0; i < n; ++i)
                    for (size_t k = 0; k <= i; ++k) lt[k * n + i] = h.factor[i * n + k];
                d.lt.upload(lt, s);
                if (coarse_mode == SF_COARSE_INVERSE) {
                    // inverse = L^-T L^-1 in f64 on the device (symmetric by construction): the
                    // columns of L^-1, then every entry, each in a fixed order
                    DeviceBuffer<double> linv;
                    std::vector<double> identity(n * n, 0.0);
                    for (size_t i = 0; i < n; ++i) identity[i * n + i] = 1.0;
                    linv.upload(identity, s);
                    d.inverse.resize(n * n);
                    for (size_t k = 0; k < n; ++k) {
                        hipLaunchKernelGGL(factor_inverse_finish_row_kernel, dim3(grid_for(k + 1, 256)), dim3(256), 0, s, (uint32_t)n,
                                           (uint32_t)k, (const double *)d.lt.ptr, linv.ptr);
                        if (k + 1 < n)
                            hipLaunchKernelGGL(factor_inverse_update_kernel, dim3(grid_for(k + 1, 256), (unsigned)(n - k - 1)), dim3(256), 0,
                                               s, (uint32_t)n, (uint32_t)k, (const double *)d.lt.ptr, linv.ptr);
                    }
                    hipLaunchKernelGGL(factor_inverse_entries_kernel, dim3(grid_for(n * n, 256)), dim3(256), 0, s, (uint32_t)n,
                                       (const double *)linv.ptr, d.inverse.ptr);
                    if (n <= COARSE_SMALL) raise_coarse_small<V>();
                    SF_HIP(hipGetLastError());
                    SF_HIP(hipStreamSynchronize(s));  // linv is freed on return
                }
                const size_t need = (n * (n + 1) / 2 + 2 * n) * sizeof(double);
                int dev = 0, limit = 0;
                SF_HIP(hipGetDevice(&dev));
                SF_HIP(hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
                if (need <= (size_t)limit) {
                    std::vector<double> packed;
                    packed.reserve(n * (n + 1) / 2);
                    for (size_t k = 0; k < n; ++k)
                        for (size_t i = k; i < n; ++i) packed.push_back(h.factor[i * n + k]);
                    d.packed.upload(packed, s);
                    d.lds_bytes = need;
                    if (need > 65536) {
                        SF_HIP(hipFuncSetAttribute((const void *)coarse_solve_lds_kernel<V>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
                        SF_HIP(hipFuncSetAttribute((const void *)coarse_solve_lds_kernel<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
                    }
                }
            }
            for (auto *buf : {&d.b, &d.x, &d.d0, &d.d1, &d.r0, &d.r1, &d.res, &d.px}) buf->resize(d.n);
            SF_HIP(hipStreamSynchronize(s));
            clock.lap(l == 0 ? "hierarchy level 0 upload" : "hierarchy coarse level upload");
        }
        in_.resize(host.empty() ? 0 : host[0].a.dofs());
        out_.resize(in_.count);
        SF_HIP(hipStreamSynchronize(s));
        tune(s);
        clock.lap("hierarchy kernel tuning");
        compute_stats();
    }

    ~Hierarchy() override {
        for (auto &g : graphs_) drop_graph(g);
    }

    void apply(const double *r, double *z, hipStream_t s) override {
        if (lv_.empty()) return;
        if (use_graph && !Profiler::get().any_on()) {
            // The captured launches hold the hierarchy's own buffers, which never move after
            // construction, and r and z, hence the key. An executable graph may be launched on
            // any stream; each replay records the entry's own event, so eviction and teardown
            // wait on that event and never on a caller's stream, which may be gone by then.
            for (auto &g : graphs_)
                if (g.r == r && g.z == z) {
                    SF_HIP(hipGraphLaunch(g.exec, s));
                    SF_HIP(hipEventRecord(g.done, s));
                    return;
                }
            if (graphs_.size() >= 8) {  // a bounded cache: the oldest goes once its last replay ended
                drop_graph(graphs_.front());
                graphs_.erase(graphs_.begin());
            }
            Graph g{r, z, nullptr, nullptr};
            SF_HIP(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) {
                (void)hipEventDestroy(g.done);
                SF_HIP(e);
            }
            try {
                launch_apply(r, z, s);
            } catch (...) {
                (void)hipStreamEndCapture(s, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                (void)hipEventDestroy(g.done);
                throw;
            }
            e = hipStreamEndCapture(s, &graph);
            if (e == hipSuccess) e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            if (graph) (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                (void)hipEventDestroy(g.done);
                SF_HIP(e);
            }
            graphs_.push_back(g);
            SF_HIP(hipGraphLaunch(g.exec, s));
            SF_HIP(hipEventRecord(g.done, s));
            return;
        }
        launch_apply(r, z, s);
    }

    static void drop_graph(Graph &g) {
        if (g.done) (void)hipEventSynchronize(g.done);
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.done) (void)hipEventDestroy(g.done);
        g.exec = nullptr;
        g.done = nullptr;
    }

    void launch_apply(const double *r, double *z, hipStream_t s) {
        if constexpr (std::is_same<V, double>::value) {
            cycle(0, r, z, s);
        } else {
            LevelDev<S, V> &d = lv_[0];
            const size_t n = d.n;
            if (d.has_p) {
                // the rounding in, fused with the first start; the widening out, fused with the
                // last step
                const unsigned g = grid_for(d.rows, 256);
                if (d.rb == 3)
                    hipLaunchKernelGGL((convert_start_kernel<3, S, V>), dim3(g), dim3(256), 0, s, (uint32_t)d.rows, r, in_.ptr,
                                       d.dinv.ptr, (V)d.theta, d.d0.ptr, out_.ptr);
                else
                    hipLaunchKernelGGL((convert_start_kernel<6, S, V>), dim3(g), dim3(256), 0, s, (uint32_t)d.rows, r, in_.ptr,
                                       d.dinv.ptr, (V)d.theta, d.d0.ptr, out_.ptr);
                cycle(0, in_.ptr, out_.ptr, s, true, z);
            } else {
                hipLaunchKernelGGL((convert_kernel<V, double>), dim3(grid_for(n, 256)), dim3(256), 0, s, n, r, in_.ptr);
                cycle(0, in_.ptr, out_.ptr, s);
                hipLaunchKernelGGL((convert_kernel<double, V>), dim3(grid_for(n, 256)), dim3(256), 0, s, n, (const V *)out_.ptr, z);
            }
        }
        SF_HIP(hipGetLastError());
    }

    void apply_batch(const double *r, double *z, uint32_t columns, hipStream_t s) override {
        if (lv_.empty() || columns == 0) return;
        if (columns > MAX_COLUMNS) throw Error(SF_ERR_INVALID_ARGUMENT, "more columns than a batch carries");
        prepare_batch(columns, s);
        if constexpr (std::is_same<V, double>::value) {
            cycle_batch(0, r, z, columns, s);
        } else {
            const size_t n = lv_[0].n * (size_t)columns;
            hipLaunchKernelGGL((convert_kernel<V, double>), dim3(grid_for(n, 256)), dim3(256), 0, s, n, r, min_.ptr);
            cycle_batch(0, min_.ptr, mout_.ptr, columns, s);
            hipLaunchKernelGGL((convert_kernel<double, V>), dim3(grid_for(n, 256)), dim3(256), 0, s, n, (const V *)mout_.ptr, z);
        }
        SF_HIP(hipGetLastError());
    }

    int tuned_depth(size_t level, int kind) const override {
        const LevelDev<S, V> &d = lv_.at(level);
        return kind == 0 ? d.a.batch_cheb : d.a.batch_spmv;
    }

    void smooth_level(size_t level, const double *b, double *x, hipStream_t s) override {
        LevelDev<S, V> &d = lv_.at(level);
        if constexpr (std::is_same<V, double>::value) {
            smooth(d, b, x, nullptr, s);
        } else {
            hipLaunchKernelGGL((convert_kernel<V, double>), dim3(grid_for(d.n, 256)), dim3(256), 0, s, d.n, b, d.b.ptr);
            smooth(d, d.b.ptr, d.x.ptr, nullptr, s);
            hipLaunchKernelGGL((convert_kernel<double, V>), dim3(grid_for(d.n, 256)), dim3(256), 0, s, d.n, (const V *)d.x.ptr, x);
        }
        SF_HIP(hipGetLastError());
    }

    void coarse_solve(const double *b, double *x, hipStream_t s) override {
        LevelDev<S, V> &d = lv_.back();
        if (!d.has_factor) throw Error(SF_ERR_MG_COARSEST_NOT_FACTORED, "the coarsest level has no direct factorization", (int64_t)d.n);
        if (d.lds_bytes)
            hipLaunchKernelGGL(coarse_solve_lds_kernel<double>, dim3(1), dim3(WAVE), d.lds_bytes, s, (uint32_t)d.n, d.packed.ptr, b, x, 1u);
        else
            hipLaunchKernelGGL(coarse_solve_kernel<double>, dim3(1), dim3(WAVE), 0, s, (uint32_t)d.n, d.lt.ptr, b, x, 1u);
        SF_HIP(hipGetLastError());
    }

   private:
    template <int RB, int BATCH>
    static void cheb_step_launch(LevelDev<S, V> &d, const V *d_in, const V *r_in, V keep, V stp, V *d_out, V *r_out, V *x,
                                 V *into, hipStream_t s, double *widen) {
        hipLaunchKernelGGL((cheb_step_kernel<RB, BATCH, S, V>), dim3(grid_for(d.a.slices, 4)), dim3(256), 0, s, d.a.view(),
                           d.dinv.ptr, d_in, r_in, keep, stp, d_out, r_out, x, into, widen);
    }
    template <int RB>
    static void cheb_step_rb(LevelDev<S, V> &d, int batch, const V *d_in, const V *r_in, V keep, V stp, V *d_out, V *r_out,
                             V *x, V *into, hipStream_t s, double *widen) {
        switch (batch) {
            case 1: cheb_step_launch<RB, 1>(d, d_in, r_in, keep, stp, d_out, r_out, x, into, s, widen); break;
            case 2: cheb_step_launch<RB, 2>(d, d_in, r_in, keep, stp, d_out, r_out, x, into, s, widen); break;
            case 8: cheb_step_launch<RB, 8>(d, d_in, r_in, keep, stp, d_out, r_out, x, into, s, widen); break;
            default: cheb_step_launch<RB, 4>(d, d_in, r_in, keep, stp, d_out, r_out, x, into, s, widen); break;
        }
    }
    static void launch_cheb_step(LevelDev<S, V> &d, int batch, const V *d_in, const V *r_in, V keep, V stp, V *d_out, V *r_out,
                                 V *x, V *into, hipStream_t s, double *widen = nullptr) {
        if (d.rb == 3) cheb_step_rb<3>(d, batch, d_in, r_in, keep, stp, d_out, r_out, x, into, s, widen);
        else cheb_step_rb<6>(d, batch, d_in, r_in, keep, stp, d_out, r_out, x, into, s, widen);
    }

    // ---- several columns ----

    template <int RB, int BATCH, int CPL>
    static void cheb_batch_launch(LevelDev<S, V> &d, uint32_t M, const V *d_in, const V *r_in, V keep, V stp, V *d_out,
                                  V *r_out, V *x, V *into, hipStream_t s, bool node) {
        if (node) {
            const size_t threads = (size_t)d.a.slices * Shape<RB>::NPW * WAVE;
            hipLaunchKernelGGL((cheb_step_node_kernel<RB, BATCH, S, V>), dim3(grid_for(threads, 256)), dim3(256), 0, s, d.a.batch_view(), M,
                               d.dinv.ptr, d_in, r_in, keep, stp, d_out, r_out, x, into);
            return;
        }
        const int lg = columns_log2((M + CPL - 1) / CPL);
        const size_t threads = (size_t)d.a.slices * Shape<RB>::NPW << lg;
        hipLaunchKernelGGL((cheb_step_batch_kernel<RB, BATCH, S, V, CPL>), dim3(grid_for(threads, 256)), dim3(256), 0, s, d.a.batch_view(),
                           lg, M, d.dinv.ptr, d_in, r_in, keep, stp, d_out, r_out, x, into);
    }
    template <int RB, int CPL>
    static void cheb_batch_depth(LevelDev<S, V> &d, int depth, uint32_t M, const V *d_in, const V *r_in, V keep, V stp,
                                 V *d_out, V *r_out, V *x, V *into, hipStream_t s, bool node) {
        switch (depth) {
            case 1: cheb_batch_launch<RB, 1, CPL>(d, M, d_in, r_in, keep, stp, d_out, r_out, x, into, s, node); break;
            case 2: cheb_batch_launch<RB, 2, CPL>(d, M, d_in, r_in, keep, stp, d_out, r_out, x, into, s, node); break;
            default: cheb_batch_launch<RB, 4, CPL>(d, M, d_in, r_in, keep, stp, d_out, r_out, x, into, s, node); break;
        }
    }
    template <int RB>
    static void cheb_batch_rb(LevelDev<S, V> &d, int batch, uint32_t M, const V *d_in, const V *r_in, V keep, V stp,
                              V *d_out, V *r_out, V *x, V *into, hipStream_t s) {
        const int depth = form_depth(batch);
        int cpl = form_columns(batch);
        if (M % (uint32_t)cpl) cpl = 1;
        if (cpl == 8) cheb_batch_depth<RB, 8>(d, depth, M, d_in, r_in, keep, stp, d_out, r_out, x, into, s, false);
        else if (cpl == 4) cheb_batch_depth<RB, 4>(d, depth, M, d_in, r_in, keep, stp, d_out, r_out, x, into, s, false);
        else if (cpl == 2) cheb_batch_depth<RB, 2>(d, depth, M, d_in, r_in, keep, stp, d_out, r_out, x, into, s, false);
        else cheb_batch_depth<RB, 1>(d, depth, M, d_in, r_in, keep, stp, d_out, r_out, x, into, s, node_form(batch));
    }
    static void launch_cheb_batch(LevelDev<S, V> &d, int batch, uint32_t M, const V *d_in, const V *r_in, V keep, V stp,
                                  V *d_out, V *r_out, V *x, V *into, hipStream_t s) {
        if (d.a.slices == 0) return;
        if (d.rb == 3) cheb_batch_rb<3>(d, batch, M, d_in, r_in, keep, stp, d_out, r_out, x, into, s);
        else cheb_batch_rb<6>(d, batch, M, d_in, r_in, keep, stp, d_out, r_out, x, into, s);
    }

    // Sizes the column work vectors for M columns and tunes the batched kernels for M's group.
    void prepare_batch(uint32_t M, hipStream_t s) {
        const int lg = columns_log2(M);
        if (lv_[0].mb.count < lv_[0].n * (size_t)M) SF_HIP(hipStreamSynchronize(s));  // buffers in use are freed below
        for (size_t l = 0; l < lv_.size(); ++l) {
            LevelDev<S, V> &d = lv_[l];
            const size_t need = d.n * (size_t)M;
            if (d.mb.count < need)
                for (auto *buf : {&d.mb, &d.mx, &d.md0, &d.md1, &d.mr0, &d.mr1, &d.mres, &d.mpx}) buf->resize(need);
        }
        const size_t top = lv_[0].n * (size_t)M;
        if (!std::is_same<V, double>::value && min_.count < top) {
            min_.resize(top);
            mout_.resize(top);
        }
        // the memo's class: the column count's group, the matrix's shape class, the vector precision
        const std::string group = std::to_string(lg) + " ", vsize = "|" + std::to_string(sizeof(V));
        for (size_t l = 0; l < lv_.size(); ++l) {
            LevelDev<S, V> &d = lv_[l];
            if (d.mm_depth[lg][0]) continue;
            const size_t need = d.n * (size_t)M;
            V *vs[8] = {d.mb.ptr, d.mx.ptr, d.md0.ptr, d.md1.ptr, d.mr0.ptr, d.mr1.ptr, d.mres.ptr, d.mpx.ptr};
            for (V *v : vs) SF_HIP(hipMemsetAsync(v, 0, need * sizeof(V), s));
            d.mm_depth[lg][0] = tune_batch(s, [&](int batch) {
                launch_cheb_batch(d, batch, M, d.md0.ptr, d.mb.ptr, (V)0, (V)0, d.md1.ptr, d.mr0.ptr, d.mx.ptr, nullptr, s);
            }, spmm_candidates(d.rb, d.rb, sizeof(S), M, false), "", "batch step " + group + shape_class(d.a) + vsize);
            d.mm_depth[lg][1] = tune_batch(s, [&](int batch) {
                launch_spmm<S, V>(d.a, M, d.mx.ptr, d.mres.ptr, d.mb.ptr, EPI_RESIDUAL, s, batch);
            }, spmm_candidates(d.rb, d.rb, sizeof(S), M, true, staged_candidate(d.a, M, sizeof(V))), "",
               "batch residual " + group + shape_class(d.a) + vsize);
            if (d.has_p) {
                LevelDev<S, V> &c = lv_[l + 1];
                const size_t cneed = c.n * (size_t)M;
                SF_HIP(hipMemsetAsync(c.mb.ptr, 0, cneed * sizeof(V), s));
                SF_HIP(hipMemsetAsync(c.mx.ptr, 0, cneed * sizeof(V), s));
                d.mm_depth[lg][2] = tune_batch(s, [&](int batch) {
                    launch_spmm<S, V>(d.r, M, d.mres.ptr, c.mb.ptr, nullptr, EPI_STORE, s, batch);
                }, spmm_candidates(d.r.rb, d.r.cb, sizeof(S), M, true, staged_candidate(d.r, M, sizeof(V))), "",
                   "batch restriction " + group + shape_class(d.r) + vsize);
                d.mm_depth[lg][3] = tune_batch(s, [&](int batch) {
                    launch_spmm<S, V>(d.p, M, c.mx.ptr, d.mpx.ptr, nullptr, EPI_ADD, s, batch);
                }, spmm_candidates(d.p.rb, d.p.cb, sizeof(S), M, true, staged_candidate(d.p, M, sizeof(V))), "",
                   "batch prolongation " + group + shape_class(d.p) + vsize);
            }
        }
        SF_HIP(hipStreamSynchronize(s));
    }

    void smooth_batch(LevelDev<S, V> &d, const V *b, V *x, V *sum_into, uint32_t M, hipStream_t s) {
        const int lg = columns_log2(M);
        const unsigned g = grid_for(d.rows * (size_t)M, 256);
        if (d.rb == 3)
            hipLaunchKernelGGL((cheb_start_batch_kernel<3, S, V>), dim3(g), dim3(256), 0, s, (uint32_t)d.rows, M, d.dinv.ptr, b, (V)d.theta, d.md0.ptr, x);
        else
            hipLaunchKernelGGL((cheb_start_batch_kernel<6, S, V>), dim3(g), dim3(256), 0, s, (uint32_t)d.rows, M, d.dinv.ptr, b, (V)d.theta, d.md0.ptr, x);
        V *dd[2] = {d.md0.ptr, d.md1.ptr};
        V *rr[2] = {d.mr0.ptr, d.mr1.ptr};
        const size_t steps = d.steps.size();
        // the finest level's steps are timed for the batched roofline: the matrix and block inverses
        // once per launch, d_in, r_in, d_out, r_out and x read or written for every column (x both),
        // and the cycle's x read and written by the last step of a post-smoothing
        int token = -1;
        if (&d == &lv_[0] && steps) {
            const double vec = (double)d.n * M * sizeof(V);
            const double each = (double)d.a.algorithmic_matrix_bytes() + (double)d.rows * d.rb * d.rb * sizeof(S) + 6 * vec;
            token = Profiler::get().begin(3, s, each * (double)steps + (sum_into ? 2 * vec : 0.0), steps);
        }
        for (size_t k = 1; k <= steps; ++k) {
            const V *d_in = dd[(k - 1) & 1];
            const V *r_in = k == 1 ? b : rr[(k - 1) & 1];
            const V keep = (V)d.steps[k - 1].first, stp = (V)d.steps[k - 1].second;
            launch_cheb_batch(d, d.mm_depth[lg][0], M, d_in, r_in, keep, stp, dd[k & 1], rr[k & 1], x, k == steps ? sum_into : nullptr, s);
        }
        Profiler::get().end(3, token, s);
        if (steps == 0 && sum_into) {
            const size_t n = d.n * (size_t)M;
            hipLaunchKernelGGL((add_kernel<V>), dim3(grid_for(n, 256)), dim3(256), 0, s, n, sum_into, (const V *)x);
        }
    }

    // cycle() for M columns.
    void cycle_batch(size_t l, const V *b, V *x, uint32_t M, hipStream_t s) {
        LevelDev<S, V> &d = lv_[l];
        const int lg = columns_log2(M);
        if (!d.has_p) {
            if (d.has_factor) {
                if (d.inverse.ptr)
                    hipLaunchKernelGGL(coarse_inverse_batch_kernel<V>, dim3(grid_for(d.n, coarse_rows(M))), dim3(256), 0, s, (uint32_t)d.n, M,
                                       coarse_rows(M), d.inverse.ptr, b, x);
                else if (d.lds_bytes)
                    hipLaunchKernelGGL(coarse_solve_lds_kernel<V>, dim3(M), dim3(WAVE), d.lds_bytes, s, (uint32_t)d.n, d.packed.ptr, b, x, M);
                else
                    hipLaunchKernelGGL(coarse_solve_kernel<V>, dim3(M), dim3(WAVE), 0, s, (uint32_t)d.n, d.lt.ptr, b, x, M);
            } else {
                smooth_batch(d, b, x, nullptr, M, s);
            }
            return;
        }
        LevelDev<S, V> &c = lv_[l + 1];
        smooth_batch(d, b, x, nullptr, M, s);
        launch_spmm<S, V>(d.a, M, x, d.mres.ptr, b, EPI_RESIDUAL, s, d.mm_depth[lg][1]);
        launch_spmm<S, V>(d.r, M, d.mres.ptr, c.mb.ptr, nullptr, EPI_STORE, s, d.mm_depth[lg][2]);
        cycle_batch(l + 1, c.mb.ptr, c.mx.ptr, M, s);
        launch_spmm<S, V>(d.p, M, c.mx.ptr, x, nullptr, EPI_ADD, s, d.mm_depth[lg][3]);
        launch_spmm<S, V>(d.a, M, x, d.mres.ptr, b, EPI_RESIDUAL, s, d.mm_depth[lg][1]);
        smooth_batch(d, d.mres.ptr, d.mpx.ptr, x, M, s);
    }

    template <int RPW>
    static void launch_cheb_rows_rpw(LevelDev<S, V> &d, const V *d_in, const V *r_in, V keep, V stp, V *d_out, V *r_out, V *x,
                                     V *into, hipStream_t s, double *widen) {
        const dim3 grid((d.a_rows.rows + RPW - 1) / RPW), block(WAVE);
        if (d.rb == 3)
            hipLaunchKernelGGL((cheb_rows_kernel<3, S, V, RPW>), grid, block, 0, s, d.a_rows.view(), d.dinv.ptr, d_in, r_in, keep, stp,
                               d_out, r_out, x, into, widen);
        else
            hipLaunchKernelGGL((cheb_rows_kernel<6, S, V, RPW>), grid, block, 0, s, d.a_rows.view(), d.dinv.ptr, d_in, r_in, keep, stp,
                               d_out, r_out, x, into, widen);
    }
    static void launch_cheb_rows(LevelDev<S, V> &d, int rpw, const V *d_in, const V *r_in, V keep, V stp, V *d_out, V *r_out, V *x,
                                 V *into, hipStream_t s, double *widen = nullptr) {
        if (rpw == 4) launch_cheb_rows_rpw<4>(d, d_in, r_in, keep, stp, d_out, r_out, x, into, s, widen);
        else if (rpw == 2) launch_cheb_rows_rpw<2>(d, d_in, r_in, keep, stp, d_out, r_out, x, into, s, widen);
        else launch_cheb_rows_rpw<1>(d, d_in, r_in, keep, stp, d_out, r_out, x, into, s, widen);
    }

    // Picks each level's batch depths, kernel forms and start fusions by timing its own kernels
    // on its own matrices. Every choice gives the same bits.
    void tune(hipStream_t s) {
        for (size_t l = 0; l < lv_.size(); ++l) {
            LevelDev<S, V> &d = lv_[l];
            const std::string at = "L" + std::to_string(l) + (sizeof(V) == 4 ? " f32 " : " f64 ");
            const std::string vs_ = "|" + std::to_string(sizeof(V)), a_class = shape_class(d.a) + vs_;
            V *vs[8] = {d.b.ptr, d.x.ptr, d.d0.ptr, d.d1.ptr, d.r0.ptr, d.r1.ptr, d.res.ptr, d.px.ptr};
            for (V *v : vs) SF_HIP(hipMemsetAsync(v, 0, d.n * sizeof(V), s));
            d.a.batch_cheb = tune_batch(s, [&](int batch) {
                launch_cheb_step(d, batch, d.d0.ptr, d.b.ptr, (V)0, (V)0, d.d1.ptr, d.r0.ptr, d.x.ptr, nullptr, s);
            }, {1, 2, 4, 8}, at + "step depth", "step " + a_class);
            d.a.batch_spmv = tune_batch(s, [&](int batch) { launch_spmv<S, V>(d.a, d.x.ptr, d.res.ptr, d.b.ptr, EPI_RESIDUAL, s, batch); },
                                        {1, 2, 4, 8}, at + "residual depth", "residual " + a_class);
            if (d.a_rows.rows) {
                // each kernel against its row-per-wavefront forms with 1, 2 or 4 rows per wavefront
                // (all the same bits); 0 is the sliced kernel
                d.cheb_rows = fastest(s, [&](int rpw) {
                    if (!rpw) launch_cheb_step(d, d.a.batch_cheb, d.d0.ptr, d.b.ptr, (V)0, (V)0, d.d1.ptr, d.r0.ptr, d.x.ptr, nullptr, s);
                    else launch_cheb_rows(d, rpw, d.d0.ptr, d.b.ptr, (V)0, (V)0, d.d1.ptr, d.r0.ptr, d.x.ptr, nullptr, s);
                }, {0, 1, 2, 4}, at + "step rows per wave", "step rows " + a_class);
                d.residual_rows = fastest(s, [&](int rpw) {
                    if (!rpw) launch_spmv<S, V>(d.a, d.x.ptr, d.res.ptr, d.b.ptr, EPI_RESIDUAL, s);
                    else launch_long_rows<S, V>(d.a_rows, d.x.ptr, d.res.ptr, d.b.ptr, EPI_RESIDUAL, s, {nullptr, (V)0, nullptr, nullptr}, rpw);
                }, {0, 1, 2, 4}, at + "residual rows per wave", "residual rows " + a_class);
            }
            if (d.has_p) {
                LevelDev<S, V> &c = lv_[l + 1];
                SF_HIP(hipMemsetAsync(c.b.ptr, 0, c.n * sizeof(V), s));
                SF_HIP(hipMemsetAsync(c.x.ptr, 0, c.n * sizeof(V), s));
                d.r.batch_spmv = tune_batch(s, [&](int batch) { launch_spmv<S, V>(d.r, d.res.ptr, c.b.ptr, nullptr, EPI_STORE, s, batch); },
                                            {1, 2, 4, 8}, at + "restriction depth", "restriction " + shape_class(d.r) + vs_);
                // the sliced product at its best depth (0) against 1, 2 or 4 rows per wavefront
                d.restrict_rows = fastest(s, [&](int rpw) {
                    if (!rpw) launch_spmv<S, V>(d.r, d.res.ptr, c.b.ptr, nullptr, EPI_STORE, s);
                    else launch_long_rows<S, V>(d.r_rows, d.res.ptr, c.b.ptr, nullptr, EPI_STORE, s, {nullptr, (V)0, nullptr, nullptr}, rpw);
                }, {0, 1, 2, 4}, at + "restriction rows per wave", "restriction rows " + shape_class(d.r) + vs_);
                d.p.batch_spmv = tune_batch(s, [&](int batch) { launch_spmv<S, V>(d.p, c.x.ptr, d.px.ptr, nullptr, EPI_ADD, s, batch); },
                                            {1, 2, 4, 8}, at + "prolongation depth", "prolongation " + shape_class(d.p) + vs_);
            }
        }
        SF_HIP(hipStreamSynchronize(s));
    }

    // The Chebyshev start z = D^-1 b / theta, d = z, x = z (cycle.rs:578-590).
    static void launch_start(LevelDev<S, V> &d, const V *b, V *x, hipStream_t s) {
        const unsigned g = grid_for(d.n, 256);
        if (d.rb == 3)
            hipLaunchKernelGGL((cheb_start_kernel<3, S, V>), dim3(g), dim3(256), 0, s, (uint32_t)d.rows, d.dinv.ptr, b, (V)d.theta, d.d0.ptr, x);
        else
            hipLaunchKernelGGL((cheb_start_kernel<6, S, V>), dim3(g), dim3(256), 0, s, (uint32_t)d.rows, d.dinv.ptr, b, (V)d.theta, d.d0.ptr, x);
    }

    // c.b = P^T d.res, with the coarser level's first start in the epilogue when `fuse`.
    static void restrict_to(LevelDev<S, V> &d, LevelDev<S, V> &c, hipStream_t s, bool fuse) {
        const Start<S, V> st = fuse ? start_of(c, c.x.ptr) : Start<S, V>{nullptr, (V)0, nullptr, nullptr};
        if (d.restrict_rows) launch_long_rows<S, V>(d.r_rows, d.res.ptr, c.b.ptr, nullptr, EPI_STORE, s, st, d.restrict_rows);
        else launch_spmv<S, V>(d.r, d.res.ptr, c.b.ptr, nullptr, EPI_STORE, s, 0, st);
    }

    // cycle.rs:570-651: start, then degree - 1 fused steps; `sum_into` takes the final add.
    // `started`: the start already ran, fused into the kernel that produced b. `widen`: the last
    // step also writes the sum into f64 (the V-cycle's widened output).
    void smooth(LevelDev<S, V> &d, const V *b, V *x, V *sum_into, hipStream_t s, bool finest = false, bool started = false,
                double *widen = nullptr) {
        if (!started) launch_start(d, b, x, s);
        V *dd[2] = {d.d0.ptr, d.d1.ptr};
        V *rr[2] = {d.r0.ptr, d.r1.ptr};
        const size_t steps = d.steps.size();
        // the finest level's steps run back to back, timed as one bracket of `steps` launches;
        // each step's algorithmic bytes: matrix, block inverse, six vector passes, and the final
        // add (and widening) in the last
        int token = -1;
        if (finest && steps) {
            const double vec = (double)d.n * sizeof(V);
            const double each = (double)d.a.algorithmic_matrix_bytes() + (double)d.rows * d.rb * d.rb * sizeof(S) + 6 * vec;
            const double last = (sum_into ? 2 * vec : 0.0) + (sum_into && widen ? (double)d.n * 8 : 0.0);
            token = Profiler::get().begin(0, s, each * (double)steps + last, steps);
        }
        for (size_t k = 1; k <= steps; ++k) {
            const V *d_in = dd[(k - 1) & 1];
            const V *r_in = k == 1 ? b : rr[(k - 1) & 1];
            V *d_out = dd[k & 1];
            V *r_out = rr[k & 1];
            const V keep = (V)d.steps[k - 1].first, stp = (V)d.steps[k - 1].second;
            V *into = k == steps ? sum_into : nullptr;
            double *wide = into ? widen : nullptr;
            if (d.cheb_rows) launch_cheb_rows(d, d.cheb_rows, d_in, r_in, keep, stp, d_out, r_out, x, into, s, wide);
            else launch_cheb_step(d, d.a.batch_cheb, d_in, r_in, keep, stp, d_out, r_out, x, into, s, wide);
        }
        Profiler::get().end(0, token, s);
        if (steps == 0 && sum_into) {
            hipLaunchKernelGGL((add_kernel<V>), dim3(grid_for(d.n, 256)), dim3(256), 0, s, d.n, sum_into, (const V *)x);
            if (widen)
                hipLaunchKernelGGL((convert_kernel<double, V>), dim3(grid_for(d.n, 256)), dim3(256), 0, s, d.n, (const V *)sum_into, widen);
        }
    }

    // res = b - A x (cycle.rs:548-564) with the level's faster kernel, optionally with the start of
    // the smoothing that follows
    static void residual(LevelDev<S, V> &d, const V *x, const V *b, hipStream_t s, Start<S, V> st = {nullptr, (V)0, nullptr, nullptr}) {
        if (d.residual_rows) launch_long_rows<S, V>(d.a_rows, x, d.res.ptr, b, EPI_RESIDUAL, s, st, d.residual_rows);
        else launch_spmv<S, V>(d.a, x, d.res.ptr, b, EPI_RESIDUAL, s, 0, st);
    }
    // Whether a level's cycle begins with a smoothing (and so with a start that can be fused).
    static bool smooths_first(const LevelDev<S, V> &d) { return d.has_p || !d.has_factor; }
    static Start<S, V> start_of(LevelDev<S, V> &d, V *x) { return {d.dinv.ptr, (V)d.theta, d.d0.ptr, x}; }

    // cycle.rs:461-539. `started`: this level's first start ran fused upstream; `widen`: the cycle's
    // output is also written to f64 there (finest level only).
    void cycle(size_t l, const V *b, V *x, hipStream_t s, bool started = false, double *widen = nullptr) {
        LevelDev<S, V> &d = lv_[l];
        if (!d.has_p) {
            if (d.has_factor) {
                if (d.inverse.ptr && d.n <= COARSE_SMALL)
                    launch_coarse_small<V>((uint32_t)d.n, d.inverse.ptr, b, x, s);
                else if (d.inverse.ptr)
                    hipLaunchKernelGGL(coarse_inverse_kernel<V>, dim3(grid_for(d.n, 256)), dim3(256), 0, s, (uint32_t)d.n, d.inverse.ptr, b, x);
                else if (d.lds_bytes)
                    hipLaunchKernelGGL(coarse_solve_lds_kernel<V>, dim3(1), dim3(WAVE), d.lds_bytes, s, (uint32_t)d.n, d.packed.ptr, b, x, 1u);
                else
                    hipLaunchKernelGGL(coarse_solve_kernel<V>, dim3(1), dim3(WAVE), 0, s, (uint32_t)d.n, d.lt.ptr, b, x, 1u);
            } else {
                smooth(d, b, x, nullptr, s, false, started);
            }
            return;
        }
        LevelDev<S, V> &c = lv_[l + 1];
        const bool finest = l == 0;
        const double res_bytes = (double)d.a.algorithmic_matrix_bytes() + 3.0 * d.n * sizeof(V);
        smooth(d, b, x, nullptr, s, finest, started);
        int token = finest ? Profiler::get().begin(2, s, res_bytes) : -1;
        residual(d, x, b, s);
        Profiler::get().end(2, token, s);
        // restriction, with the coarser level's first start when its cycle smooths first
        const bool fuse = smooths_first(c);
        restrict_to(d, c, s, fuse);
        cycle(l + 1, c.b.ptr, c.x.ptr, s, fuse);
        launch_spmv<S, V>(d.p, c.x.ptr, x, nullptr, EPI_ADD, s);
        // the residual for the post-smoothing, with its start
        token = finest ? Profiler::get().begin(2, s, res_bytes + (double)d.rows * d.rb * d.rb * sizeof(S) + 2.0 * d.n * sizeof(V)) : -1;
        residual(d, x, b, s, start_of(d, d.px.ptr));
        Profiler::get().end(2, token, s);
        smooth(d, d.res.ptr, d.px.ptr, x, s, finest, true, widen);
    }

    void compute_stats() {
        // MultigridStats::of_levels, cycle.rs:676-699, with stored values in S
        const double finest = host.empty() ? 1.0 : (double)(host[0].a.blocks() * host[0].a.block_len());
        size_t stored = 0, memory = 0;
        for (const auto &h : host) {
            const size_t values = h.a.blocks() * h.a.block_len();
            stored += values;
            memory += h.a.starts.size() * 8 + h.a.columns.size() * 4 + values * sizeof(S);
            memory += h.inverse_diagonal.size() * sizeof(S);
            if (h.has_prolongator)
                memory += h.prolongator.starts.size() * 8 + h.prolongator.columns.size() * 4 +
                          h.prolongator.blocks() * h.prolongator.block_len() * sizeof(S);
            if (h.has_factor) memory += h.factor.size() * sizeof(double);
        }
        stats.levels = host.size();
        stats.coarse_dofs = host.empty() ? 0 : host.back().a.dofs();
        stats.operator_complexity = (double)stored / finest;
        stats.memory_bytes = memory;
        // algorithmic bytes of one V-cycle
        const double sv = sizeof(V), ss = sizeof(S);
        double bytes = 0;
        const double k = (double)degree;
        for (const auto &d : lv_) {
            const double n = (double)d.n;
            const double abytes = (double)d.a.algorithmic_matrix_bytes();
            const double dbytes = (double)d.rows * d.rb * d.rb * ss;
            if (!d.has_p) {
                if (d.has_factor) bytes += n * n * 8 + 2 * n * sv;
                else bytes += dbytes + 3 * n * sv + (k - 1) * (abytes + dbytes + 6 * n * sv);
                continue;
            }
            const double smooth_bytes = (dbytes + 3 * n * sv) + (k - 1) * (abytes + dbytes + 6 * n * sv);
            bytes += 2 * smooth_bytes + n * sv;                                   // + final add of x
            bytes += 2 * (abytes + 3 * n * sv);                                   // two residuals
            bytes += (double)d.r.algorithmic_matrix_bytes() + n * sv + (double)d.r.dofs() * sv;  // restrict
            bytes += (double)d.p.algorithmic_matrix_bytes() + (double)d.r.dofs() * sv + 2 * n * sv;  // prolong
        }
        cycle_bytes = bytes;
    }

    std::vector<LevelDev<S, V>> lv_;
    DeviceBuffer<V> in_, out_;
    DeviceBuffer<V> min_, mout_;  // the finest level's columns in V, for the single precision
};

}  // namespace

std::unique_ptr<DeviceMultigrid> make_device_multigrid(std::vector<HostLevel> levels, size_t degree, int precision,
                                                       int coarse_mode, hipStream_t s, const DeviceSliced<double> *finest) {
    if (degree < 1) throw Error(SF_ERR_INVALID_ARGUMENT, "the smoother degree must be at least one");
    if (coarse_mode != SF_COARSE_REFERENCE && coarse_mode != SF_COARSE_INVERSE)
        throw Error(SF_ERR_INVALID_ARGUMENT, "unknown coarse solve");
    switch (precision) {
        case SF_PRECISION_DOUBLE:
            return std::make_unique<Hierarchy<double, double>>(std::move(levels), degree, precision, coarse_mode, s, finest);
        case SF_PRECISION_SINGLE_STORAGE:
            return std::make_unique<Hierarchy<float, double>>(std::move(levels), degree, precision, coarse_mode, s, finest);
        case SF_PRECISION_SINGLE:
            return std::make_unique<Hierarchy<float, float>>(std::move(levels), degree, precision, coarse_mode, s, finest);
        default:
            throw Error(SF_ERR_INVALID_ARGUMENT, "unknown precision");
    }
}

// ---- the system matrix ------------------------------------------------------------------------

void DeviceSystem::upload_from_host(hipStream_t s) {
    values.upload(host.values, s);
    build_from_device_values(s);
}

void DeviceSystem::build_from_device_values(hipStream_t s) {
    StageClock clock;
    SlicedPlan plan = plan_sliced(host.starts.data(), host.columns.data(), host.rows, (int)host.rb, (int)host.cb);
    clock.lap("system layout plan");
    build_sliced(sliced, plan, host.cols, values.ptr, s);
    SF_HIP(hipStreamSynchronize(s));
    clock.lap("system layout build");
    tune_spmv<double, double>(sliced, EPI_STORE, s);
    clock.lap("system product tuning");
}

// ---- the conjugate gradient -------------------------------------------------------------------

std::vector<double> jacobi_inverse(const std::shared_ptr<Context> &ctx, const DeviceSystem &a) {
    // Jacobi::new, preconditioner.rs:53-66
    const auto &h = a.host;
    const size_t n = a.size;
    std::vector<double> inv(n);
    // A study keeps its values on the device: its diagonal blocks come to the host.
    std::vector<double> diagonal;
    std::vector<char> has_diagonal;
    if (h.values.empty() && h.blocks() > 0) {
        if (h.rb != h.cb) throw Error(SF_ERR_INTERNAL, "a system matrix of non-square blocks");
        std::vector<int64_t> slots(h.rows, 0);
        has_diagonal.assign(h.rows, 0);
        for (size_t br = 0; br < h.rows; ++br) {
            auto first = h.columns.begin() + h.starts[br], last = h.columns.begin() + h.starts[br + 1];
            auto it = std::lower_bound(first, last, (u32)br);
            if (it != last && *it == (u32)br) {
                slots[br] = it - h.columns.begin();
                has_diagonal[br] = 1;
            }
        }
        diagonal.resize(h.rows * h.block_len());
        DeviceValues(ctx, a.values.ptr, h.blocks(), h.block_len()).blocks(slots, diagonal.data());
    }
    for (size_t row = 0; row < n; ++row) {
        const size_t brow = row / h.rb, p = row % h.rb;
        double diag = 0.0;
        if (!diagonal.empty()) {
            if (has_diagonal[brow]) diag = diagonal[brow * h.block_len() + p * h.cb + p];
        } else {
            for (int64_t at = h.starts[brow]; at < h.starts[brow + 1]; ++at)
                if (h.columns[(size_t)at] * h.cb <= row && row < (h.columns[(size_t)at] + 1) * h.cb)
                    diag = h.values[(size_t)at * h.rb * h.cb + p * h.cb + (row - h.columns[(size_t)at] * h.cb)];
        }
        if (std::isnan(diag) || diag <= 0.0 || !std::isfinite(diag))
            throw Error(SF_ERR_JACOBI_DIAGONAL, "a stiffness matrix needs a positive diagonal", (int64_t)row, 0, diag);
        inv[row] = 1.0 / diag;
    }
    return inv;
}

DevicePcg::DevicePcg(std::shared_ptr<Context> ctx, const DeviceSystem &a, DeviceMultigrid *mg, const std::vector<double> &rhs,
                     const sf_pcg_options &options, const sf_preconditioner *user)
    : ctx_(std::move(ctx)), a_(a), mg_(mg), n_(a.size), dot_order_(options.dot_order), tolerance_(options.tolerance) {
    if (user) {
        if (!user->apply) throw Error(SF_ERR_INVALID_ARGUMENT, "the preconditioner has no apply callback");
        if (user->where != SF_APPLY_HOST && user->where != SF_APPLY_DEVICE)
            throw Error(SF_ERR_INVALID_ARGUMENT, "unknown preconditioner placement", user->where);
        if (mg_) throw Error(SF_ERR_INVALID_ARGUMENT, "a solve takes one preconditioner");
        user_ = *user;
    }
    if (rhs.size() != n_) throw Error(SF_ERR_INVALID_ARGUMENT, "the right-hand side does not match the matrix");
    if (mg_ && mg_->dofs() != n_) throw Error(SF_ERR_INVALID_ARGUMENT, "the multigrid does not match the matrix");
    hipStream_t s = ctx_->stream;
    max_iterations_ = options.max_iterations ? options.max_iterations : 10 * (u64)n_ + 1000;
    for (auto *buf : {&b_, &x_, &r_, &z_, &p_, &ap_, &tmp_, &tmp2_}) buf->resize(std::max<size_t>(n_, 1));
    b_.upload(rhs.data(), n_, s);
    scalars_.resize(1);
    partials_.resize(DOT_BLOCKS);
    partials2_.resize(DOT_BLOCKS);
    norm_.resize(1);
    SF_HIP(hipHostMalloc((void **)&pinned_, sizeof(PcgScalars), hipHostMallocCoherent));
    std::memset((void *)pinned_, 0, sizeof(PcgScalars));
    SF_HIP(hipHostGetDevicePointer((void **)&mirror_, pinned_, 0));
    if (user_.apply && user_.where == SF_APPLY_HOST) {
        SF_HIP(hipHostMalloc((void **)&host_r_, std::max<size_t>(n_, 1) * 8, hipHostMallocDefault));
        SF_HIP(hipHostMalloc((void **)&host_z_, std::max<size_t>(n_, 1) * 8, hipHostMallocDefault));
    }
    if (!mg_ && !user_.apply) jacobi_.upload(jacobi_inverse(ctx_, a_), s);
    initialise();
}

void DevicePcg::initialise() {
    // pcg.rs:135-164: r = b, x = 0, then the first direction from the preconditioned residual
    hipStream_t s = ctx_->stream;
    SF_HIP(hipMemcpyAsync(r_.ptr, b_.ptr, n_ * 8, hipMemcpyDeviceToDevice, s));
    SF_HIP(hipMemsetAsync(x_.ptr, 0, n_ * 8, s));
    SF_HIP(hipMemsetAsync(z_.ptr, 0, n_ * 8, s));
    SF_HIP(hipMemsetAsync(p_.ptr, 0, n_ * 8, s));
    SF_HIP(hipMemsetAsync(scalars_.ptr, 0, sizeof(PcgScalars), s));
    iterations_ = 0;
    relative_ = 0;
    floor_ = false;
    rz_slot_ = 0;
    status_ = PcgStatus::Active;
    dot(b_.ptr, b_.ptr, norm_.ptr);
    double bb = 0;
    norm_.download(&bb, 1, s);
    SF_HIP(hipStreamSynchronize(s));
    rhs_norm_ = std::sqrt(bb);
    if (rhs_norm_ == 0.0) {
        status_ = PcgStatus::Converged;
    } else {
        reset_direction();
        relative_ = 1.0;
    }
    SF_HIP(hipStreamSynchronize(s));
}

void DevicePcg::restart(const double *new_rhs) {
    SF_HIP(hipStreamSynchronize(ctx_->stream));
    if (new_rhs) b_.upload(new_rhs, n_, ctx_->stream);
    initialise();
}

DevicePcg::~DevicePcg() {
    (void)hipStreamSynchronize(ctx_->stream);
    if (pinned_) (void)hipHostFree(pinned_);
    if (host_r_) (void)hipHostFree(host_r_);
    if (host_z_) (void)hipHostFree(host_z_);
}

void DevicePcg::dot(const double *a, const double *b, double *result, int final_kind, PcgScalars *mirror, uint32_t tag) {
    if (dot_order_ == SF_DOT_REFERENCE) {
        hipLaunchKernelGGL((dot_serial_kernel<0>), dim3(1), dim3(64), 0, ctx_->stream, n_, a, b, final_kind, result, scalars_.ptr, mirror,
                           tag);
    } else {
        hipLaunchKernelGGL((dot_partials_kernel<0>), dim3(DOT_BLOCKS), dim3(DOT_THREADS), 0, ctx_->stream, n_, a, b, partials_.ptr);
        hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(DOT_THREADS), 0, ctx_->stream, DOT_BLOCKS, (const double *)partials_.ptr,
                           final_kind, result, scalars_.ptr, mirror, tag);
    }
    SF_HIP(hipGetLastError());
}

void DevicePcg::sum(const double *a, double *result) {
    if (dot_order_ == SF_DOT_REFERENCE) {
        hipLaunchKernelGGL((dot_serial_kernel<1>), dim3(1), dim3(64), 0, ctx_->stream, n_, a, a, (int)FINAL_STORE, result, scalars_.ptr,
                           (PcgScalars *)nullptr, 0u);
    } else {
        hipLaunchKernelGGL((dot_partials_kernel<1>), dim3(DOT_BLOCKS), dim3(DOT_THREADS), 0, ctx_->stream, n_, a, a, partials_.ptr);
        hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(DOT_THREADS), 0, ctx_->stream, DOT_BLOCKS, (const double *)partials_.ptr,
                           (int)FINAL_STORE, result, scalars_.ptr, (PcgScalars *)nullptr, 0u);
    }
    SF_HIP(hipGetLastError());
}

void DevicePcg::precondition(const double *r, double *z) {
    hipStream_t s = ctx_->stream;
    if (user_.apply) {
        int32_t rc = 0;
        if (user_.where == SF_APPLY_DEVICE) {
            rc = user_.apply(r, z, (uint64_t)n_, (void *)s, user_.user);
        } else {
            // the previous z's copy to the device is ordered before this copy on the stream, so
            // the staging is free once this synchronisation returns
            SF_HIP(hipMemcpyAsync(host_r_, r, n_ * 8, hipMemcpyDeviceToHost, s));
            SF_HIP(hipStreamSynchronize(s));
            rc = user_.apply(host_r_, host_z_, (uint64_t)n_, (void *)s, user_.user);
            if (rc == 0) SF_HIP(hipMemcpyAsync(z, host_z_, n_ * 8, hipMemcpyHostToDevice, s));
        }
        if (rc != 0) throw Error(SF_ERR_PRECONDITIONER, "the caller's preconditioner failed", rc);
        return;
    }
    if (mg_) mg_->apply(r, z, s);
    else hipLaunchKernelGGL(jacobi_kernel, dim3(grid_for(n_, 256)), dim3(256), 0, s, n_, jacobi_.ptr, r, z);
    SF_HIP(hipGetLastError());
}

void DevicePcg::reset_direction() {
    // pcg.rs:277-282
    precondition(r_.ptr, z_.ptr);
    SF_HIP(hipMemcpyAsync(p_.ptr, z_.ptr, n_ * 8, hipMemcpyDeviceToDevice, ctx_->stream));
    dot(r_.ptr, z_.ptr, rz_slot_ ? &scalars_.ptr->rz_next : &scalars_.ptr->rz);
}

// Waits until the host copy of the scalars carries `tag`. The copy is written by a kernel already
// enqueued; should it not arrive within a second of spinning, the stream is synchronised and the
// copy checked once more, so a lost write shows as an error rather than a hang.
void DevicePcg::wait_for_tag(uint32_t tag) {
    const auto start = std::chrono::steady_clock::now();
    for (uint64_t spins = 0;; ++spins) {
        if (__atomic_load_n(&pinned_->tag, __ATOMIC_ACQUIRE) == tag) return;
        if ((spins & 1023) == 1023 && std::chrono::steady_clock::now() - start > std::chrono::seconds(1)) break;
    }
    SF_HIP(hipStreamSynchronize(ctx_->stream));
    if (__atomic_load_n(&pinned_->tag, __ATOMIC_ACQUIRE) != tag)
        throw Error(SF_ERR_INTERNAL, "the solver's scalars did not reach the host");
}

PcgScalars DevicePcg::read_scalars() {
    SF_HIP(hipMemcpyAsync(pinned_, scalars_.ptr, sizeof(PcgScalars), hipMemcpyDeviceToHost, ctx_->stream));
    SF_HIP(hipStreamSynchronize(ctx_->stream));
    return *pinned_;
}

void DevicePcg::iterate() {
    // pcg.rs:216-275. The branch taken when the recursive residual is above the tolerance (the
    // preconditioner, r.z and the new direction) is enqueued before the residual is read back, so
    // the readback overlaps the V-cycle. When the other branch is taken, that work is discarded:
    // it changes only z, p and rz, which the true-residual branch resets or never reads again.
    hipStream_t s = ctx_->stream;
    PcgScalars *sc = scalars_.ptr;
    const int token = Profiler::get().begin(1, s, (double)a_.sliced.algorithmic_matrix_bytes() + 2.0 * n_ * 8);
    launch_spmv<double, double>(a_.sliced, p_.ptr, ap_.ptr, nullptr, EPI_STORE, s);
    Profiler::get().end(1, token, s);
    const uint32_t tag = ++readbacks_;
    if (dot_order_ == SF_DOT_TREE) {
        // the same reductions in the same order, their final sums formed by the kernels that use
        // them: p.Ap's by the update, which also forms r.r's partials, and r.z's by the direction
        hipLaunchKernelGGL((dot_partials_kernel<0>), dim3(DOT_BLOCKS), dim3(DOT_THREADS), 0, s, n_, (const double *)p_.ptr,
                           (const double *)ap_.ptr, partials_.ptr);
        hipLaunchKernelGGL(pcg_update_fused_kernel, dim3(DOT_BLOCKS), dim3(DOT_THREADS), 0, s, n_, sc, rz_slot_,
                           (const double *)partials_.ptr, (const double *)p_.ptr, (const double *)ap_.ptr, x_.ptr, r_.ptr,
                           partials2_.ptr);
        // r.r's final sum also copies the scalars into host memory, tagged for this readback
        hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(DOT_THREADS), 0, s, DOT_BLOCKS, (const double *)partials2_.ptr,
                           (int)FINAL_STORE, &sc->rr, sc, mirror_, tag);
        precondition(r_.ptr, z_.ptr);
        hipLaunchKernelGGL((dot_partials_kernel<0>), dim3(DOT_BLOCKS), dim3(DOT_THREADS), 0, s, n_, (const double *)r_.ptr,
                           (const double *)z_.ptr, partials_.ptr);
        hipLaunchKernelGGL(pcg_direction_fused_kernel, dim3(DOT_BLOCKS), dim3(DOT_THREADS), 0, s, n_, sc, rz_slot_,
                           (const double *)partials_.ptr, (const double *)z_.ptr, p_.ptr);
        rz_slot_ ^= 1u;
    } else {
        dot(p_.ptr, ap_.ptr, &sc->curvature);
        hipLaunchKernelGGL(pcg_update_kernel, dim3(grid_for(n_, 256)), dim3(256), 0, s, n_, sc, p_.ptr, ap_.ptr, x_.ptr, r_.ptr);
        // r.r, whose final kernel also copies the scalars into host memory, tagged for this readback
        dot(r_.ptr, r_.ptr, &sc->rr, FINAL_STORE, mirror_, tag);
        precondition(r_.ptr, z_.ptr);
        dot(r_.ptr, z_.ptr, nullptr, FINAL_RZ);
        hipLaunchKernelGGL(pcg_direction_kernel, dim3(grid_for(n_, 256)), dim3(256), 0, s, n_, (const PcgScalars *)sc, z_.ptr,
                           p_.ptr);
    }
    SF_HIP(hipGetLastError());
    wait_for_tag(tag);
    const PcgScalars h = *pinned_;
    if (!h.valid) {
        SF_HIP(hipStreamSynchronize(s));
        status_ = PcgStatus::NotPositiveDefinite;
        failure_ = Error(SF_ERR_PCG_NOT_POSITIVE_DEFINITE, "the system is not positive definite", (int64_t)iterations_, 0, h.curvature);
        throw failure_;
    }
    iterations_ += 1;
    relative_ = std::sqrt(h.rr) / rhs_norm_;
    if (relative_ <= tolerance_) {
        launch_spmv<double, double>(a_.sliced, x_.ptr, tmp_.ptr, b_.ptr, EPI_RESIDUAL, s);
        dot(tmp_.ptr, tmp_.ptr, &sc->true_rr);
        const dim3 grid(grid_for(a_.sliced.slices, 4)), block(256);
        if (a_.rb == 3) hipLaunchKernelGGL((rounding_bound_kernel<3, 3>), grid, block, 0, s, a_.sliced.view(), x_.ptr, tmp2_.ptr);
        else hipLaunchKernelGGL((rounding_bound_kernel<1, 1>), grid, block, 0, s, a_.sliced.view(), x_.ptr, tmp2_.ptr);
        sum(tmp2_.ptr, &sc->bound2);
        const PcgScalars t = read_scalars();
        relative_ = std::sqrt(t.true_rr) / rhs_norm_;
        const double floor = std::sqrt(t.bound2) / rhs_norm_;
        floor_ = relative_ > tolerance_ && relative_ <= floor;
        if (relative_ <= tolerance_ || floor_) {
            status_ = PcgStatus::Converged;
            return;
        }
        // rounding made the recursive residual over-optimistic: restart from the true one
        SF_HIP(hipMemcpyAsync(r_.ptr, tmp_.ptr, n_ * 8, hipMemcpyDeviceToDevice, s));
        reset_direction();
    }
    if (iterations_ >= max_iterations_) {
        SF_HIP(hipStreamSynchronize(s));
        status_ = PcgStatus::NotConverged;
        failure_ = Error(SF_ERR_PCG_NOT_CONVERGED, "the solver did not converge", (int64_t)iterations_, 0, relative_);
        throw failure_;
    }
}

int DevicePcg::step(u64 max, const Cancel &cancel) {
    // pcg.rs:168-189
    switch (status_) {
        case PcgStatus::Active:
            break;
        case PcgStatus::Converged:
            return SF_CONVERGED;
        case PcgStatus::Cancelled:
            return SF_CANCELLED;
        default:
            throw failure_;
    }
    for (u64 i = 0; i < max; ++i) {
        if (cancel()) {
            SF_HIP(hipStreamSynchronize(ctx_->stream));
            status_ = PcgStatus::Cancelled;
            return SF_CANCELLED;
        }
        iterate();
        if (status_ == PcgStatus::Converged) {
            SF_HIP(hipStreamSynchronize(ctx_->stream));
            return SF_CONVERGED;
        }
    }
    SF_HIP(hipStreamSynchronize(ctx_->stream));
    return SF_RUNNING;
}

void DevicePcg::estimate(double *out) const {
    x_.download(out, n_, ctx_->stream);
    SF_HIP(hipStreamSynchronize(ctx_->stream));
}

sf_progress DevicePcg::progress() const {
    sf_progress p{};
    p.iteration = iterations_;
    p.relative_residual = relative_;
    p.at_rounding_floor = floor_ ? 1 : 0;
    p.outcome = status_ == PcgStatus::Converged ? SF_CONVERGED : status_ == PcgStatus::Cancelled ? SF_CANCELLED : SF_RUNNING;
    return p;
}

}  // namespace sf
