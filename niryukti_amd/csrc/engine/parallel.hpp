// Minimal fork-join loop over std::thread for host passes whose per-index work
// is independent (row sums, elementwise maps).  Results never depend on the
// thread count because each index is computed by exactly one thread.
#pragma once
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <exception>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <utility>
#include <vector>

namespace nyk {

inline int host_threads() {
    static int n = [] {
        unsigned h = std::thread::hardware_concurrency();
        return int(std::clamp(h ? h : 1u, 1u, 16u));
    }();
    return n;
}

template <class F>
void parallel_for(int64_t count, bool enable, F &&body) {
    int threads = enable ? std::min<int64_t>(host_threads(), std::max<int64_t>(1, count / 4096)) : 1;
    if (threads <= 1) {
        for (int64_t i = 0; i < count; ++i) body(i);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(threads);
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            int64_t lo = count * t / threads, hi = count * (t + 1) / threads;
            for (int64_t i = lo; i < hi; ++i) body(i);
        });
    for (auto &th : pool) th.join();
}

// Coarse-grained tasks (e.g. one LP each): up to host_threads() workers pull
// task indices from a shared counter.
template <class F>
void parallel_tasks(int64_t count, F &&body) {
    const int threads = int(std::min<int64_t>(host_threads(), count));
    if (threads <= 1) {
        for (int64_t i = 0; i < count; ++i) body(i);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    pool.reserve(threads);
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (int64_t i; (i = next.fetch_add(1)) < count;) body(i);
        });
    for (auto &th : pool) th.join();
}

// One task on a thread of its own beside the caller's work (the batched
// engine certifies a round's stopped LPs this way while the next round runs).
// wait() joins the task and rethrows what it threw; the task must touch only
// data the caller leaves alone until wait() returns.
class Beside {
  public:
    Beside() = default;
    Beside(const Beside &) = delete;
    Beside &operator=(const Beside &) = delete;
    ~Beside() {
        if (thread_.joinable()) thread_.join();
    }
    template <class F>
    void start(F &&task) {
        if (thread_.joinable()) throw std::logic_error("a task is already running beside");
        error_ = nullptr;
        thread_ = std::thread([this, f = std::forward<F>(task)]() mutable {
            try {
                f();
            } catch (...) {
                error_ = std::current_exception();
            }
        });
    }
    bool running() const { return thread_.joinable(); }
    void wait() {
        if (thread_.joinable()) thread_.join();
        if (error_) {
            std::exception_ptr e = error_;
            error_ = nullptr;
            std::rethrow_exception(e);
        }
    }

  private:
    std::thread thread_;
    std::exception_ptr error_;
};

// Items handed in from any thread and taken, in order, by one consumer (the
// batched engine's submissions: any thread submits, the scheduling thread
// files them by ticket).  push() gives each item the next sequence number
// (0, 1, ...) under the same lock that appends it, so drain() returns the
// items in sequence order and item n always lands at position n of a list
// the consumer appends them to.
template <class T>
class Inbox {
  public:
    // `number(item, n)` runs under the lock with the item's sequence number.
    template <class Number>
    int64_t push(T item, Number &&number) {
        std::lock_guard<std::mutex> lock(mu_);
        const int64_t n = next_++;
        number(item, n);
        items_.push_back(std::move(item));
        return n;
    }
    // Moves every item pushed so far to the end of `out`, in sequence order.
    void drain(std::vector<T> &out) {
        std::lock_guard<std::mutex> lock(mu_);
        for (auto &item : items_) out.push_back(std::move(item));
        items_.clear();
    }
    int64_t pushed() const {
        std::lock_guard<std::mutex> lock(mu_);
        return next_;
    }

  private:
    mutable std::mutex mu_;
    std::vector<T> items_;
    int64_t next_ = 0;
};

// Destroys what it is given on a thread of its own: freeing large host
// buffers returns pages to the system, which a latency-critical thread (the
// batched engine's scheduling thread) should not wait for.  The thread starts
// with the first item; the destructor destroys what is left and joins it.
template <class T>
class Reaper {
  public:
    Reaper() = default;
    Reaper(const Reaper &) = delete;
    Reaper &operator=(const Reaper &) = delete;
    ~Reaper() {
        {
            std::lock_guard<std::mutex> lock(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        if (thread_.joinable()) thread_.join();
    }
    void give(T item) {
        {
            std::lock_guard<std::mutex> lock(mu_);
            if (!thread_.joinable()) thread_ = std::thread([this] { loop(); });
            items_.push_back(std::move(item));
        }
        cv_.notify_one();
    }

  private:
    void loop() {
        std::unique_lock<std::mutex> lock(mu_);
        for (;;) {
            cv_.wait(lock, [&] { return stop_ || !items_.empty(); });
            std::vector<T> batch;
            batch.swap(items_);
            lock.unlock();
            batch.clear();  // the destructors run outside the lock
            lock.lock();
            if (stop_ && items_.empty()) return;
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::vector<T> items_;
    bool stop_ = false;
    std::thread thread_;
};

}  // namespace nyk
