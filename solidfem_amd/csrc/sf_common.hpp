// Shared definitions of the solidfem HIP library: status codes, the error type thrown inside
// the library (never across the C ABI), HIP error checks, device buffers and a small
// deterministic host thread pool.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/solidfem_hip.h"

namespace sf {

using u32 = uint32_t;
using u64 = uint64_t;
constexpr u32 FIXED = 0xffffffffu;  // a node with no free index (crates/solidfem-elastic/src/assembly.rs:15)

// An error raised inside the library. The C ABI catches it, records the message and payload,
// and returns `status`.
struct Error : std::runtime_error {
    sf_status status;
    int64_t a = 0, b = 0;  // payload: iteration/level/row/expected ...
    double value = 0;      // payload: curvature/pivot/residual ...
    Error(sf_status s, const std::string &what, int64_t pa = 0, int64_t pb = 0, double v = 0)
        : std::runtime_error(what), status(s), a(pa), b(pb), value(v) {}
};

#define SF_HIP(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            throw ::sf::Error(SF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

// A device allocation owned by the host object that created it.
template <typename T>
struct DeviceBuffer {
    T *ptr = nullptr;
    size_t count = 0;
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) { resize(n); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&o) noexcept : ptr(o.ptr), count(o.count) {
        o.ptr = nullptr;
        o.count = 0;
    }
    DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
        if (this != &o) {
            release();
            ptr = o.ptr;
            count = o.count;
            o.ptr = nullptr;
            o.count = 0;
        }
        return *this;
    }
    ~DeviceBuffer() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    void resize(size_t n) {
        release();
        if (n) SF_HIP(hipMalloc((void **)&ptr, n * sizeof(T)));
        count = n;
    }
    size_t bytes() const { return count * sizeof(T); }
    // Copies and waits: host sources are often temporaries.
    void upload(const T *host, size_t n, hipStream_t s = nullptr) {
        if (n > count) resize(n);
        if (n) {
            SF_HIP(hipMemcpyAsync(ptr, host, n * sizeof(T), hipMemcpyHostToDevice, s));
            SF_HIP(hipStreamSynchronize(s));
        }
    }
    void upload(const std::vector<T> &host, hipStream_t s = nullptr) {
        if (s) SF_HIP(hipStreamSynchronize(s));  // nothing pending may still use the old buffer
        resize(host.size());
        upload(host.data(), host.size(), s);
    }
    void download(T *host, size_t n, hipStream_t s = nullptr) const {
        if (n) SF_HIP(hipMemcpyAsync(host, ptr, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

// A host array of a trivial type whose elements start uninitialised, so that large arrays are
// first touched by the (parallel) code that fills them rather than zeroed serially.
template <typename T>
struct HostArray {
    std::unique_ptr<T[]> data_;
    size_t size_ = 0;
    HostArray() = default;
    explicit HostArray(size_t n) : data_(n ? new T[n] : nullptr), size_(n) {}
    size_t size() const { return size_; }
    T *data() { return data_.get(); }
    const T *data() const { return data_.get(); }
    T &operator[](size_t i) { return data_[i]; }
    const T &operator[](size_t i) const { return data_[i]; }
};

// Runs `body(first, last)` over [0, n) in contiguous chunks on up to `threads` threads. Each
// index's work must depend only on its own inputs, so the chunking never changes a result.
void parallel_for(size_t n, size_t min_chunk, const std::function<void(size_t, size_t)> &body);
size_t host_threads();

// A cancel poll: the caller's callback, or never.
struct Cancel {
    sf_cancel_fn fn = nullptr;
    void *user = nullptr;
    bool operator()() const { return fn ? fn(user) != 0 : false; }
};

// SOLIDFEM_SETUP_LOG: the wall time of each stage of a study's construction, on stderr.
struct StageClock {
    std::chrono::steady_clock::time_point last = std::chrono::steady_clock::now();
    static bool on() {
        static const bool log = std::getenv("SOLIDFEM_SETUP_LOG") != nullptr;
        return log;
    }
    void lap(const char *what) {
        if (!on()) return;
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[build] %s %.1f ms\n", what, 1e3 * std::chrono::duration<double>(now - last).count());
        last = now;
    }
};

}  // namespace sf
