// parallel.cpp -- host thread pool (parallel.hpp).
#include "parallel.hpp"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <thread>
#include <vector>

namespace st {

int host_threads()
{
    static const int n = [] {
        const char *env = std::getenv("SUPERTRACK_THREADS");
        int v = env ? std::atoi(env) : 0;
        if (v <= 0) v = (int)std::thread::hardware_concurrency();
        if (v <= 0) v = 1;
        return v > 64 ? 64 : v;
    }();
    return n;
}

void parallel_for(int64_t n, const std::function<void(int64_t)> &fn)
{
    const int nt = (int)std::min<int64_t>(host_threads(), std::max<int64_t>(n, 1));
    if (nt <= 1) {
        for (int64_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&]() {
            for (int64_t i = next.fetch_add(1); i < n; i = next.fetch_add(1)) fn(i);
        });
    for (auto &th : pool) th.join();
}

}  // namespace st
