// emp_host_threads.h -- per-puzzle host work of a large batch spread over a
// few threads (puzzle checks, table building, the table blob's assembly).
#ifndef EMP_HOST_THREADS_H
#define EMP_HOST_THREADS_H

#include <algorithm>
#include <cstdint>
#include <exception>
#include <system_error>
#include <thread>
#include <vector>

namespace emp {

constexpr uint32_t kHostThreads = 16;

// Calls fn(i) for every i in [0, n), on one thread below `serial_below`
// items and on up to kHostThreads threads (i dealt round-robin) above.  Once
// every thread has finished, the exception of the lowest i that raised one
// is rethrown, so the error reported does not depend on the thread count.
template <typename Fn>
void for_each_item(uint32_t n, Fn fn, uint32_t serial_below = 64) {
    std::vector<std::exception_ptr> errors(n);
    auto run = [&](uint32_t first, uint32_t step) {
        for (uint32_t i = first; i < n; i += step) {
            try {
                fn(i);
            } catch (...) {
                errors[i] = std::current_exception();
            }
        }
    };
    const uint32_t hw = std::max(1u, std::thread::hardware_concurrency());
    const uint32_t n_threads = n < serial_below ? 1u : std::min(kHostThreads, hw);
    // Shares whose thread could not be started (a process thread limit) are
    // run by the calling thread.
    std::vector<std::thread> pool;
    pool.reserve(n_threads);
    uint32_t started = 1;
    try {
        for (; started < n_threads; ++started) pool.emplace_back(run, started, n_threads);
    } catch (const std::system_error &) {
    }
    run(0, n_threads);
    for (uint32_t t = started; t < n_threads; ++t) run(t, n_threads);
    for (std::thread &th : pool) th.join();
    for (const std::exception_ptr &e : errors)
        if (e) std::rethrow_exception(e);
}

}  // namespace emp

#endif
