// Host-side data parallelism for the construction and planning steps: a
// parallel_for over index ranges on plain std::threads.  Work split this way
// must not depend on the split (each index writes only its own outputs), so
// results are identical for any thread count.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <thread>
#include <vector>

namespace mlp {

// Threads for host work: MLP_HOST_THREADS if set, else the hardware threads,
// at most 16 (the host share of one GPU on the target machines).
inline unsigned host_threads() {
    if (const char* e = std::getenv("MLP_HOST_THREADS")) {
        const int v = std::atoi(e);
        if (v >= 1) return unsigned(std::min(v, 256));
    }
    const unsigned h = std::thread::hardware_concurrency();
    return std::max(1u, std::min(h ? h : 1u, 16u));
}

// f(begin, end) over [0, n) in contiguous chunks of at least `grain` indices.
template <class F>
void parallel_for(uint64_t n, uint64_t grain, F&& f) {
    const unsigned T = host_threads();
    const uint64_t chunks = std::min<uint64_t>(T, (n + grain - 1) / std::max<uint64_t>(grain, 1));
    if (chunks <= 1) {
        if (n) f(uint64_t(0), n);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(chunks - 1);
    const uint64_t step = (n + chunks - 1) / chunks;
    for (uint64_t c = 1; c < chunks; c++) {
        const uint64_t b = c * step, e = std::min(n, b + step);
        if (b < e) pool.emplace_back([&f, b, e] { f(b, e); });
    }
    f(uint64_t(0), std::min(n, step));
    for (auto& t : pool) t.join();
}

}  // namespace mlp
