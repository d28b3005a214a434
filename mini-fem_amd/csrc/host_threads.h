// Host threads for the one-time setup: the count (MF_PLAN_THREADS overrides it, at most
// 64; default the hardware's, at most 16) and a fork-join over thread indices. Work is
// split so that the results never depend on the thread count.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <exception>
#include <thread>
#include <vector>

namespace mf {

inline int host_threads()
{
    if (const char *e = std::getenv("MF_PLAN_THREADS")) {
        const int n = std::atoi(e);
        if (n >= 1) return std::min(n, 64);
    }
    const unsigned h = std::thread::hardware_concurrency();
    return std::max(1, std::min(16, (int)h));
}

// Runs f(t) for t = 0 .. T-1, t = 0 on the calling thread, and joins every thread before
// it returns or rethrows the first exception any of them raised. A thread that cannot be
// started runs its share on the caller.
template <class F>
void fork_join(int T, F f)
{
    std::vector<std::thread> pool;
    std::vector<std::exception_ptr> err((size_t)std::max(T, 1));
    pool.reserve((size_t)std::max(T - 1, 0));
    auto run = [&](int t) {
        try {
            f(t);
        } catch (...) {
            err[(size_t)t] = std::current_exception();
        }
    };
    for (int t = 1; t < T; t++) {
        try {
            pool.emplace_back(run, t);
        } catch (...) {
            run(t);
        }
    }
    run(0);
    for (auto &th : pool) th.join();
    for (auto &e : err)
        if (e) std::rethrow_exception(e);
}

// [lo, hi) of part t of n items cut into T contiguous parts.
inline int64_t part_lo(int64_t n, int T, int t) { return n * t / T; }

}  // namespace mf
