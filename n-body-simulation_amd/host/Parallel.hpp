// Parallel.hpp -- a small fork-join helper for the host-side formatting and
// parsing work (output writers, CSV input).  Work items are handed out by an
// atomic counter; the first exception thrown by any item is re-thrown on the
// calling thread after every worker has joined.
#ifndef NBODY_HOST_PARALLEL_HPP
#define NBODY_HOST_PARALLEL_HPP

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstdlib>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace host_parallel {

// Worker threads for host formatting and parsing: NB_HOST_THREADS when set,
// else the hardware thread count, at most 16 (the host share of one GPU).
inline unsigned defaultThreads() {
    if (const char *s = std::getenv("NB_HOST_THREADS")) {
        const long v = std::strtol(s, nullptr, 10);
        if (v >= 1) return (unsigned)std::min(v, 256L);
    }
    const unsigned hw = std::thread::hardware_concurrency();
    return std::max(1u, std::min(hw == 0 ? 1u : hw, 16u));
}

// Calls f(i) for every i in [0, count) on up to `threads` threads.
template <class F>
void parallelFor(size_t count, unsigned threads, F &&f) {
    if (threads <= 1 || count <= 1) {
        for (size_t i = 0; i < count; ++i) f(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::exception_ptr error;
    std::mutex m;
    auto work = [&] {
        for (;;) {
            const size_t i = next.fetch_add(1);
            if (i >= count) return;
            try {
                f(i);
            } catch (...) {
                std::lock_guard<std::mutex> g(m);
                if (!error) error = std::current_exception();
            }
        }
    };
    const size_t nthreads = std::min<size_t>(threads, count);
    std::vector<std::thread> pool;
    pool.reserve(nthreads - 1);
    for (size_t t = 1; t < nthreads; ++t) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    if (error) std::rethrow_exception(error);
}

}  // namespace host_parallel

#endif
