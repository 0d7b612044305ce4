// host_parallel.h - fork-join helpers for the host preparation code (BVH
// build, tangent frames). Every parallel pass in this library writes
// disjoint outputs, or combines per-thread partials in a fixed order, so its
// results do not depend on the thread count.
#pragma once

#include <algorithm>
#include <cstdlib>
#include <exception>
#include <thread>
#include <vector>

namespace ptcore {

// Worker threads for host preparation: PT_HOST_THREADS when set (1 runs on
// the calling thread alone), else the hardware's count capped at 16.
inline unsigned host_threads() {
    if (const char* v = std::getenv("PT_HOST_THREADS")) {
        long t = std::strtol(v, nullptr, 10);
        if (t >= 1) return (unsigned)std::min(t, 256L);
    }
    unsigned hw = std::thread::hardware_concurrency();
    return std::max(1u, std::min(hw, 16u));
}

// Run f(t) for t in [0, threads) on as many threads, the caller taking t = 0;
// the first exception any of them threw is rethrown after all have joined.
template <class F>
void run_parallel(unsigned threads, F&& f) {
    if (threads <= 1) {
        f(0u);
        return;
    }
    std::vector<std::exception_ptr> errors(threads);
    auto guarded = [&](unsigned t) {
        try {
            f(t);
        } catch (...) {
            errors[t] = std::current_exception();
        }
    };
    std::vector<std::thread> pool;
    pool.reserve(threads - 1);
    try {
        for (unsigned t = 1; t < threads; ++t) pool.emplace_back(guarded, t);
    } catch (...) {
        for (std::thread& th : pool) th.join();
        throw;
    }
    guarded(0u);
    for (std::thread& th : pool) th.join();
    for (std::exception_ptr& e : errors)
        if (e) std::rethrow_exception(e);
}

// f(begin, end) over `threads` contiguous slices of [0, n)
template <class F>
void run_slices(size_t n, unsigned threads, F&& f) {
    run_parallel(threads, [&](unsigned t) { f(n * t / threads, n * (t + 1) / threads); });
}

}  // namespace ptcore
