// The adaptive driver's host worker pool (ramses-cpp_amd/csrc/rh_pool.h)
// under ThreadSanitizer, compiled and run by tests/test_host_sanitizers.py
// with the host compiler; no GPU.
//
// Two host threads, each holding a pool reference as a live adaptive run
// does, run parallel passes at the same time (the pool takes their passes in
// turn); the pool is then released, which joins its workers, and acquired
// again.  Every pass must visit each index exactly once.
#include <atomic>
#include <cstdio>
#include <thread>
#include <vector>

#include "rh_pool.h"

static std::atomic<int> failures{0};

static void passes(int seed) {
    for (int rep = 0; rep < 40; ++rep) {
        const int n = 5000 + 97 * ((rep + seed) % 11);
        std::vector<int> hits(n, 0);
        rhi::parallel_for(n, [&](int i) { hits[i] += 1; }, 1024);
        for (int i = 0; i < n; ++i)
            if (hits[i] != 1) {
                failures.fetch_add(1);
                break;
            }
    }
}

int main() {
    for (int cycle = 0; cycle < 3; ++cycle) {
        rhi::pool_acquire();
        rhi::pool_acquire();
        std::thread a([] { passes(1); });
        std::thread b([] { passes(2); });
        a.join();
        b.join();
        rhi::pool_release();
        passes(3);  // one reference left: the pool is still there
        rhi::pool_release();  // the last one: workers joined
        passes(4);  // no pool: passes run on the calling thread
    }
    std::printf("pool workers: %d, %d failures\n", rhi::host_threads() - 1, failures.load());
    return failures.load() ? 1 : 0;
}
