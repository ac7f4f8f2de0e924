// parallel.hpp -- a minimal host thread pool for the library's host-side
// preparation work (synthetic tracks, library layout).  Plain C++.
#pragma once

#include <cstdint>
#include <functional>

namespace st {

// Worker threads: SUPERTRACK_THREADS if set, else the hardware's, at most 64.
int host_threads();
// Calls fn(i) for i in [0, n) on host_threads() threads (dynamic schedule).
void parallel_for(int64_t n, const std::function<void(int64_t)> &fn);

}  // namespace st
