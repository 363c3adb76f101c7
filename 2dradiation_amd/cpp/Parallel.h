// Minimal host thread pool loop for set-up work (metric arrays).  std::thread
// keeps the library free of an OpenMP runtime.
#ifndef RT2D_PARALLEL_H
#define RT2D_PARALLEL_H
#include <algorithm>
#include <thread>
#include <vector>

template <class F>
inline void ParallelFor(size_t n, F &&f)
{
    size_t nt = std::max<size_t>(1, std::min<size_t>(std::thread::hardware_concurrency(), 32));
    if (n < 4096 || nt == 1) {
        for (size_t k = 0; k < n; k++) f(k);
        return;
    }
    std::vector<std::thread> pool;
    size_t chunk = (n + nt - 1) / nt;
    for (size_t t = 0; t < nt; t++) {
        size_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([b, e, &f] {
            for (size_t k = b; k < e; k++) f(k);
        });
    }
    for (auto &th : pool) th.join();
}
#endif
