// rh_pool.h -- the host worker pool of the adaptive driver's flag passes.
// Host-only (no HIP): the library includes it from rh_amr.cpp, and the
// thread-sanitizer test builds it with the host compiler.
#pragma once
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace rhi {

// Worker threads shared by every adaptive run of the process: a flag pass
// is a few hundred microseconds of work, comparable to starting and joining
// threads, so the passes share one pool instead of spawning their own.  The
// pool lives while some run does (pool_acquire in rh_amr_create,
// pool_release in rh_amr_destroy, which joins the workers when the last run
// goes; a pool still alive when the library is unloaded is joined then).
// run() hands out chunks [k*chunk, (k+1)*chunk) through an atomic counter;
// the caller works too and returns once every chunk is done.  Passes from
// different host threads take turns (run_mu_).
class HostPool {
  public:
    explicit HostPool(int nworkers) {
        for (int i = 0; i < nworkers; ++i) threads_.emplace_back([this] { worker(); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (std::thread& t : threads_) t.join();
    }
    void run(int nchunks, const std::function<void(int)>& fn) {
        std::lock_guard<std::mutex> one_pass(run_mu_);
        {
            std::lock_guard<std::mutex> lk(m_);
            // A worker still leaving the previous pass may take a chunk of
            // this one as soon as next_ is reset: everything it reads is set
            // first.
            fn_.store(&fn);
            nchunks_.store(nchunks);
            left_.store(nchunks);
            next_.store(0);
            ++gen_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [this] { return left_.load() == 0; });
        fn_.store(nullptr);
    }

  private:
    void work() {
        for (int k; (k = next_.fetch_add(1)) < nchunks_.load();) {
            (*fn_.load())(k);
            if (left_.fetch_sub(1) == 1) {
                std::lock_guard<std::mutex> lk(m_);
                done_.notify_all();
            }
        }
    }
    void worker() {
        unsigned long long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
                if (!fn_.load()) continue;
            }
            work();
        }
    }
    std::mutex run_mu_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    std::atomic<const std::function<void(int)>*> fn_{nullptr};
    std::atomic<int> nchunks_{0}, next_{0}, left_{0};
    unsigned long long gen_ = 0;
    bool stop_ = false;
    std::vector<std::thread> threads_;
};

inline int host_threads() {
    static const int n = [] {
        unsigned h = std::thread::hardware_concurrency();
        return (int)std::max(1u, std::min(h == 0 ? 1u : h, 16u));
    }();
    return n;
}

inline std::mutex g_pool_mu;  // guards g_pool_users and the creation / release of g_pool
inline int g_pool_users = 0;
inline std::atomic<HostPool*> g_pool{nullptr};
// Joins a pool left alive when the library is unloaded or the process ends.
struct PoolAtUnload {
    ~PoolAtUnload() { delete g_pool.exchange(nullptr); }
};
inline PoolAtUnload g_pool_at_unload;

inline void pool_acquire() {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if (g_pool_users++ == 0 && host_threads() > 1) g_pool.store(new HostPool(host_threads() - 1));
}

inline void pool_release() {
    HostPool* p = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        if (--g_pool_users == 0) p = g_pool.exchange(nullptr);
    }
    delete p;  // joins the workers
}

// Runs fn(i) for i in [0, n) on the host's cores (the flag passes touch each
// oct's own cells only, or defer their writes, so octs are independent).
// Called only by a live run, which holds a pool reference.
template <class Fn> void parallel_for(int n, Fn fn, int min_parallel = 4096) {
    HostPool* pool = g_pool.load();
    if (n < min_parallel || !pool) {
        for (int i = 0; i < n; ++i) fn(i);
        return;
    }
    const int nchunks = 4 * host_threads();
    const int chunk = (n + nchunks - 1) / nchunks;
    const std::function<void(int)> body = [&](int k) {
        const int lo = k * chunk, hi = std::min(n, lo + chunk);
        for (int i = lo; i < hi; ++i) fn(i);
    };
    pool->run(nchunks, body);
}

}  // namespace rhi
