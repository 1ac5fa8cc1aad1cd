// host_pool.hpp -- a small persistent pool of host threads for the byte copies between the
// caller's (pageable) arrays and the library's pinned staging buffers.  One memcpy thread
// cannot keep up with a PCIe DMA engine; several can.
#pragma once

#include <condition_variable>
#include <cstddef>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace magace {

class HostPool {
public:
    // the process-wide pool, made at the first call (again in a forked child); its size is
    // MAGACE_COPY_THREADS if set (1..64), else min(8, hardware threads).  It lives until the
    // process ends.
    static HostPool& get();

    explicit HostPool(int nthreads);
    ~HostPool();
    HostPool(const HostPool&) = delete;
    HostPool& operator=(const HostPool&) = delete;

    // threads taking part in run(), the caller included
    int size() const { return (int)workers_.size() + 1; }

    // calls fn(0), ..., fn(ntasks - 1) on the workers and the calling thread; returns when
    // every call has returned.  Not reentrant: one run() at a time.
    void run(int ntasks, const std::function<void(int)>& fn);

private:
    void work();
    void drain();

    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable start_, done_;
    const std::function<void(int)>* fn_ = nullptr;
    int ntasks_ = 0;
    int next_ = 0;       // next task index to hand out
    int finished_ = 0;   // tasks completed
    unsigned gen_ = 0;   // bumped by every run()
    bool stop_ = false;
};

struct CopyJob {
    void* dst;
    const void* src;
    size_t bytes;
};

// copies every job, each split into pieces so that all threads of the pool take part
void parallel_copy(HostPool& pool, const std::vector<CopyJob>& jobs);

}  // namespace magace
