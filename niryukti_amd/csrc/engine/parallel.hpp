// Minimal fork-join loop over std::thread for host passes whose per-index work
// is independent (row sums, elementwise maps).  Results never depend on the
// thread count because each index is computed by exactly one thread.
#pragma once
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <thread>
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

}  // namespace nyk
