// Host-side parallel loops of the planners.
#pragma once
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace t3 {

// Calls f(i) for 0 <= i < n on up to T3NS_PLAN_THREADS host threads (default:
// the hardware threads, at most 16), indices taken from a shared counter in
// chunks; the first exception thrown is rethrown here once every thread has
// stopped.  Work below `min_work` (the caller's estimate, in its own units)
// runs on the calling thread: starting threads would cost more.
template <typename F>
void run_blocks_threaded(int n, F&& f, double work = 1e30, double min_work = 0) {
    int nt = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (const char* e = std::getenv("T3NS_PLAN_THREADS")) nt = std::max(1, std::atoi(e));
    nt = std::min(nt, std::max(1, n / 4));
    if (work < min_work) nt = 1;
    if (nt <= 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int> next{0};
    std::atomic<bool> failed{false};
    std::exception_ptr err;
    std::mutex mu;
    auto worker = [&]() {
        try {
            for (;;) {
                const int i0 = next.fetch_add(4);
                if (i0 >= n || failed.load()) break;
                for (int i = i0; i < std::min(n, i0 + 4); ++i) f(i);
            }
        } catch (...) {
            std::lock_guard<std::mutex> lk(mu);
            if (!err) err = std::current_exception();
            failed.store(true);
        }
    };
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (int t = 1; t < nt; ++t) th.emplace_back(worker);
    worker();
    for (auto& x : th) x.join();
    if (err) std::rethrow_exception(err);
}

}  // namespace t3
