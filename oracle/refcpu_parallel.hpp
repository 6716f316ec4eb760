// refcpu: TEST INFRASTRUCTURE (the oracle) -- the row-chunk threading of
// crates/solidfem-solve/src/parallel.rs: a persistent pool that runs one closure per chunk of rows
// (WorkerPool, :82-215), chunk bounds balanced by stored values (balanced_bounds, :220-239) and
// the region helper (for_rows, :252-292). Every value a chunk writes is a function of its own row
// alone, so the bits never depend on the thread count; the sums across rows (dot products, the
// transpose product) stay serial in their callers.
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace refcpu {

// parallel.rs:28: levels and systems with fewer stored values than this run on one thread
constexpr size_t MIN_PARALLEL_VALUES = 100000;
// parallel.rs:31-38
constexpr size_t MAX_DEFAULT_THREADS = 8;
constexpr size_t MAX_THREADS = 64;

// parallel.rs:43-49: half the logical processors, between 1 and 8
inline size_t default_threads() {
    const size_t logical = std::max<size_t>(1, std::thread::hardware_concurrency());
    return std::min(std::max<size_t>(logical / 2, 1), MAX_DEFAULT_THREADS);
}

// parallel.rs:82-215: `threads` threads including the caller. run(task) calls task(0) on the
// caller and task(i) on worker i, and returns once all have finished; a failure in any of them
// is rethrown after they have all stopped.
class WorkerPool {
   public:
    explicit WorkerPool(size_t threads) {
        threads = std::min(threads, MAX_THREADS);
        for (size_t index = 1; index < threads; ++index) workers_.emplace_back([this, index] { work(index); });
    }
    ~WorkerPool() {
        {
            std::lock_guard<std::mutex> g(m_);
            shutdown_ = true;
        }
        start_.notify_all();
        for (auto &t : workers_) t.join();
    }
    size_t threads() const { return workers_.size() + 1; }

    void run(const std::function<void(size_t)> &task) {
        {
            std::lock_guard<std::mutex> g(m_);
            task_ = &task;
            region_ += 1;
            running_ = workers_.size();
            failure_ = nullptr;
        }
        start_.notify_all();
        std::exception_ptr own;
        try {
            task(0);
        } catch (...) {
            own = std::current_exception();
        }
        std::unique_lock<std::mutex> l(m_);
        finished_.wait(l, [this] { return running_ == 0; });
        task_ = nullptr;
        std::exception_ptr other = failure_;
        l.unlock();
        if (own) std::rethrow_exception(own);
        if (other) std::rethrow_exception(other);
    }

   private:
    void work(size_t index) {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(size_t)> *task;
            {
                std::unique_lock<std::mutex> l(m_);
                start_.wait(l, [&] { return shutdown_ || region_ != seen; });
                if (shutdown_) return;
                seen = region_;
                task = task_;
            }
            std::exception_ptr failure;
            try {
                (*task)(index);
            } catch (...) {
                failure = std::current_exception();
            }
            std::lock_guard<std::mutex> g(m_);
            if (failure && !failure_) failure_ = failure;
            if (--running_ == 0) finished_.notify_one();
        }
    }

    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable start_, finished_;
    const std::function<void(size_t)> *task_ = nullptr;
    uint64_t region_ = 0;
    size_t running_ = 0;
    bool shutdown_ = false;
    std::exception_ptr failure_;
};

// WorkerPool::new (parallel.rs:94-124): no pool for one thread
inline std::unique_ptr<WorkerPool> make_pool(size_t threads) {
    if (std::min(threads, MAX_THREADS) <= 1) return nullptr;
    return std::make_unique<WorkerPool>(threads);
}

// parallel.rs:220-239: at most `chunks` contiguous row ranges of about equal weight, from the
// rows' cumulative weights (cumulative[row] is the weight before `row`, the last entry the
// total); the bounds run from zero to the row count and no chunk is empty.
inline std::vector<size_t> balanced_bounds(const std::vector<size_t> &cumulative, size_t chunks) {
    const size_t rows = cumulative.empty() ? 0 : cumulative.size() - 1;
    const size_t total = cumulative.empty() ? 0 : cumulative.back();
    std::vector<size_t> bounds{0};
    for (size_t chunk = 1; chunk < chunks; ++chunk) {
        const size_t target = total * chunk / chunks;
        const size_t row = std::min<size_t>(
            (size_t)(std::lower_bound(cumulative.begin(), cumulative.end(), target) - cumulative.begin()), rows);
        if (row > bounds.back() && row < rows) bounds.push_back(row);
    }
    if (rows > 0) bounds.push_back(rows);
    return bounds;
}

// parallel.rs:252-292: work(first, last) once per chunk of `bounds`, on the pool when there is
// more than one chunk, else on the caller
template <typename F>
inline void for_rows(WorkerPool *pool, const std::vector<size_t> &bounds, const F &work) {
    const size_t rows = bounds.empty() ? 0 : bounds.back();
    if (!pool || bounds.size() <= 2) {
        work((size_t)0, rows);
        return;
    }
    pool->run([&](size_t chunk) {
        if (chunk + 1 >= bounds.size()) return;
        work(bounds[chunk], bounds[chunk + 1]);
    });
}

}  // namespace refcpu
