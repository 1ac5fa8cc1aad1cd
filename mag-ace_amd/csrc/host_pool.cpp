// host_pool.cpp -- persistent host threads for parallel staging copies.
#include "host_pool.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include <unistd.h>

namespace magace {

HostPool& HostPool::get() {
    // one pool per process: a child made by fork() inherits the object but not its threads,
    // so it builds a pool of its own (the parent's copy is left alone; its threads are not
    // the child's to join)
    static std::mutex mu;
    static HostPool* pool = nullptr;
    static pid_t owner = 0;
    std::lock_guard<std::mutex> lk(mu);
    if (!pool || owner != getpid()) {
        int n = (int)std::thread::hardware_concurrency();
        n = std::max(1, std::min(8, n));
        if (const char* s = std::getenv("MAGACE_COPY_THREADS")) {
            const int v = std::atoi(s);
            if (v >= 1 && v <= 64) n = v;
        }
        pool = new HostPool(n);
        owner = getpid();
    }
    return *pool;
}

HostPool::HostPool(int nthreads) {
    for (int t = 1; t < nthreads; ++t) workers_.emplace_back([this] { work(); });
}

HostPool::~HostPool() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    start_.notify_all();
    for (auto& w : workers_) w.join();
}

// hands out task indices until none is left; called with mu_ unlocked
void HostPool::drain() {
    for (;;) {
        int i;
        const std::function<void(int)>* fn;
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (next_ >= ntasks_) return;
            i = next_++;
            fn = fn_;
        }
        (*fn)(i);
        bool last;
        {
            std::lock_guard<std::mutex> lk(mu_);
            last = ++finished_ == ntasks_;
        }
        if (last) done_.notify_all();
    }
}

void HostPool::work() {
    unsigned seen = 0;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu_);
            start_.wait(lk, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
        }
        drain();
    }
}

void HostPool::run(int ntasks, const std::function<void(int)>& fn) {
    if (ntasks <= 0) return;
    if (workers_.empty() || ntasks == 1) {
        for (int i = 0; i < ntasks; ++i) fn(i);
        return;
    }
    {
        std::lock_guard<std::mutex> lk(mu_);
        fn_ = &fn;
        ntasks_ = ntasks;
        next_ = 0;
        finished_ = 0;
        ++gen_;
    }
    start_.notify_all();
    drain();
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return finished_ == ntasks_; });
    fn_ = nullptr;
    ntasks_ = 0;
}

void parallel_copy(HostPool& pool, const std::vector<CopyJob>& jobs) {
    size_t total = 0;
    for (const auto& j : jobs) total += j.bytes;
    if (total == 0) return;
    // pieces of at least 256 KiB, about four per thread
    const size_t piece = std::max<size_t>(size_t(256) << 10, total / (4 * (size_t)pool.size()) + 1);
    struct Piece {
        char* dst;
        const char* src;
        size_t bytes;
    };
    std::vector<Piece> pieces;
    for (const auto& j : jobs)
        for (size_t o = 0; o < j.bytes; o += piece)
            pieces.push_back({(char*)j.dst + o, (const char*)j.src + o, std::min(piece, j.bytes - o)});
    pool.run((int)pieces.size(), [&](int i) { std::memcpy(pieces[i].dst, pieces[i].src, pieces[i].bytes); });
}

}  // namespace magace
