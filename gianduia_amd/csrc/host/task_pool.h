// task_pool.h — a small fork-join thread pool for the host scene build.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace gnd {

// A small fork-join pool: tasks go to one queue; a thread waiting for its
// tasks runs queued ones meanwhile, so nested waits cannot deadlock.
class TaskPool {
public:
    explicit TaskPool(int workers) {
        for (int i = 0; i < workers; ++i) threads_.emplace_back([this] { loop(); });
    }
    ~TaskPool() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : threads_) t.join();
    }
    int size() const { return (int)threads_.size() + 1; }
    void submit(std::function<void()> f) {
        {
            std::lock_guard<std::mutex> g(mu_);
            q_.push_back(std::move(f));
            queued_.fetch_add(1, std::memory_order_release);
        }
        cv_.notify_one();
    }
    // runs one queued task; false if the queue was empty (checked without
    // the lock first, so waiting threads do not contend with submitters)
    bool runOne() {
        if (queued_.load(std::memory_order_acquire) == 0) return false;
        std::function<void()> f;
        {
            std::lock_guard<std::mutex> g(mu_);
            if (q_.empty()) return false;
            f = std::move(q_.front());
            q_.pop_front();
            queued_.fetch_sub(1, std::memory_order_relaxed);
        }
        f();
        return true;
    }
    void waitFor(const std::atomic<int>& pending) {
        while (pending.load(std::memory_order_acquire) > 0)
            if (!runOne()) std::this_thread::yield();
    }
    // fn(lo, hi) over [0, n) in about 4 pieces per thread.  The caller and up
    // to one helper task per worker pull pieces from a private counter; the
    // caller then waits only for pieces in flight (running unrelated tasks
    // there could hold up its own continuation).  A helper that starts after
    // every piece was taken touches only the shared counters.
    template <class F>
    void parallelFor(int64_t n, F&& fn) {
        if (n <= 0) return;
        const int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(n, 4 * size()));
        struct Shared {
            std::atomic<int64_t> next{0}, done{0};
        };
        auto sh = std::make_shared<Shared>();
        auto pull = [sh, n, chunks](F& f) {
            for (;;) {
                const int64_t c = sh->next.fetch_add(1, std::memory_order_relaxed);
                if (c >= chunks) return;
                f(n * c / chunks, n * (c + 1) / chunks);
                sh->done.fetch_add(1, std::memory_order_release);
            }
        };
        const int helpers = (int)std::min<int64_t>(chunks - 1, size() - 1);
        for (int h = 0; h < helpers; ++h) submit([pull, &fn]() mutable { pull(fn); });
        pull(fn);
        while (sh->done.load(std::memory_order_acquire) < chunks) std::this_thread::yield();
    }
private:
    void loop() {
        for (;;) {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [this] { return stop_ || !q_.empty(); });
                if (stop_ && q_.empty()) return;
                f = std::move(q_.front());
                q_.pop_front();
                queued_.fetch_sub(1, std::memory_order_relaxed);
            }
            f();
        }
    }
    std::vector<std::thread> threads_;
    std::deque<std::function<void()>> q_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::atomic<int> queued_{0};
    bool stop_ = false;
};

}  // namespace gnd
