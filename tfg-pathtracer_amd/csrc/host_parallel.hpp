// Host-side data parallelism for the scene-ingestion builders: fn(begin, end)
// over contiguous chunks of [0, n) on up to hostThreads() threads (the
// calling thread takes the first chunk).  Small ranges run inline.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <thread>
#include <vector>

namespace eleven {

// Threads the host builders use: hardware_concurrency, capped by
// OMP_NUM_THREADS when it is set (a job's CPU share on a shared node).
inline unsigned hostThreads() {
    unsigned t = std::max(1u, std::thread::hardware_concurrency());
    if (const char* e = std::getenv("OMP_NUM_THREADS")) {
        const long v = std::strtol(e, nullptr, 10);
        if (v > 0) t = std::min(t, (unsigned)v);
    }
    return t;
}

template <class F>
void parallelFor(size_t n, F&& fn, size_t minChunk = 16384) {
    const size_t hw = hostThreads();
    const size_t parts = std::min(hw, std::max<size_t>(1, n / minChunk));
    if (parts <= 1) {
        if (n) fn((size_t)0, n);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(parts - 1);
    for (size_t p = 1; p < parts; p++) pool.emplace_back([&fn, n, parts, p] { fn(n * p / parts, n * (p + 1) / parts); });
    fn((size_t)0, n / parts);
    for (auto& t : pool) t.join();
}

}  // namespace eleven
