// Host-side data parallelism for the construction and planning steps: a
// parallel_for over index ranges.  Work split this way must not depend on the
// split (each index writes only its own outputs), so results are identical for
// any thread count.
//
// Top-level calls run on a pool of persistent worker threads (creating threads
// per call costs more than the smaller loops themselves).  A call made from
// inside a pool task, or while another thread is using the pool, runs on
// threads of its own instead.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <memory>
#include <mutex>
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

// Allocator whose value-less construct leaves elements uninitialised, so that
// large host buffers are first written (and their pages first touched) by the
// threads that fill them.
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U> struct rebind { using other = NoInitAlloc<U>; };
    NoInitAlloc() = default;
    template <class U> NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
    template <class U> void construct(U* p) noexcept { ::new (static_cast<void*>(p)) U; }
    template <class U, class... Args> void construct(U* p, Args&&... args) {
        ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
    }
};
template <class T> using HostVec = std::vector<T, NoInitAlloc<T>>;

namespace detail {

// Persistent workers.  run(n, task) has the workers and the caller take task
// indices 0..n-1 from a shared counter and returns when all are done and no
// worker is still inside this round, so the next round can reset the counter.
class Pool {
public:
    explicit Pool(unsigned workers) {
        for (unsigned i = 0; i < workers; i++) th_.emplace_back([this] { loop(); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    std::mutex& user() { return user_; }
    // true on a worker thread, and on a caller thread while its round runs
    static bool& in_worker() {
        thread_local bool w = false;
        return w;
    }
    void run(uint64_t n, const std::function<void(uint64_t)>& task) {
        {
            std::lock_guard<std::mutex> g(m_);
            task_ = &task;
            ntask_ = n;
            next_.store(0);
            done_ = 0;
            gen_++;
        }
        cv_.notify_all();
        const uint64_t mine = drain(task, n);
        std::unique_lock<std::mutex> lk(m_);
        done_ += mine;
        done_cv_.wait(lk, [&] { return done_ == n && active_ == 0; });
        task_ = nullptr;
    }

private:
    uint64_t drain(const std::function<void(uint64_t)>& task, uint64_t n) {
        uint64_t k, c = 0;
        while ((k = next_.fetch_add(1)) < n) {
            task(k);
            c++;
        }
        return c;
    }
    void loop() {
        in_worker() = true;
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait(lk, [&] { return stop_ || (gen_ != seen && task_); });
            if (stop_) return;
            seen = gen_;
            const std::function<void(uint64_t)>* t = task_;
            const uint64_t n = ntask_;
            active_++;
            lk.unlock();
            const uint64_t mine = drain(*t, n);
            lk.lock();
            done_ += mine;
            active_--;
            if (done_ == n && active_ == 0) done_cv_.notify_all();
        }
    }
    std::vector<std::thread> th_;
    std::mutex user_, m_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(uint64_t)>* task_ = nullptr;
    uint64_t ntask_ = 0, done_ = 0, gen_ = 0;
    unsigned active_ = 0;
    std::atomic<uint64_t> next_{0};
    bool stop_ = false;
};

inline Pool& pool() {
    static Pool p(host_threads() - 1);
    return p;
}

}  // namespace detail

// f(begin, end) over [0, n) in contiguous chunks of at least `grain` indices.
template <class F>
void parallel_for(uint64_t n, uint64_t grain, F&& f) {
    const unsigned T = host_threads();
    const uint64_t chunks = std::min<uint64_t>(T, (n + grain - 1) / std::max<uint64_t>(grain, 1));
    if (chunks <= 1) {
        if (n) f(uint64_t(0), n);
        return;
    }
    const uint64_t step = (n + chunks - 1) / chunks;
    bool& inside = detail::Pool::in_worker();        // a pool worker, or the caller of a pool round
    if (!inside) {
        detail::Pool& P = detail::pool();
        std::unique_lock<std::mutex> use(P.user(), std::try_to_lock);
        if (use.owns_lock()) {
            const std::function<void(uint64_t)> task = [&](uint64_t c) {
                const uint64_t b = c * step, e = std::min(n, b + step);
                if (b < e) f(b, e);
            };
            inside = true;
            P.run(chunks, task);
            inside = false;
            return;
        }
    }
    std::vector<std::thread> pool;
    pool.reserve(chunks - 1);
    for (uint64_t c = 1; c < chunks; c++) {
        const uint64_t b = c * step, e = std::min(n, b + step);
        if (b < e) pool.emplace_back([&f, b, e] { f(b, e); });
    }
    f(uint64_t(0), std::min(n, step));
    for (auto& t : pool) t.join();
}

}  // namespace mlp
