// Host-side loop parallelism for the set-up code: parallel_for(n, fn) calls
// fn(i, worker) for every i in [0, n) on up to 16 threads (fewer for small n),
// items handed out in chunks by an atomic counter.  worker is in
// [0, parallel_workers()) and indexes per-thread scratch.  Every item is
// independent in the callers, so results do not depend on the thread count.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <thread>
#include <vector>

namespace cfdg {

inline int parallel_workers() {
    const unsigned hw = std::thread::hardware_concurrency();
    return static_cast<int>(std::max(1u, std::min(16u, hw)));
}

template <class F>
void parallel_for(std::size_t n, F&& fn, std::size_t chunk = 64) {
    const int nw = static_cast<int>(std::min<std::size_t>(static_cast<std::size_t>(parallel_workers()),
                                                          (n + chunk - 1) / std::max<std::size_t>(chunk, 1)));
    if (nw <= 1) {
        for (std::size_t i = 0; i < n; ++i) fn(i, 0);
        return;
    }
    std::atomic<std::size_t> next{0};
    const auto run = [&](int w) {
        for (;;) {
            const std::size_t b = next.fetch_add(chunk);
            if (b >= n) break;
            const std::size_t e = std::min(n, b + chunk);
            for (std::size_t i = b; i < e; ++i) fn(i, w);
        }
    };
    std::vector<std::thread> th;
    th.reserve(static_cast<std::size_t>(nw) - 1);
    for (int w = 1; w < nw; ++w) th.emplace_back(run, w);
    run(0);
    for (auto& t : th) t.join();
}

}  // namespace cfdg
