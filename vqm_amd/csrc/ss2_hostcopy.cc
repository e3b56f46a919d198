// Multi-threaded host copies (see ss2_hostcopy.h).

#include "ss2_hostcopy.h"

#include <sched.h>
#include <unistd.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

namespace vqm::ss2 {

namespace {

constexpr uint64_t kMinChunk = 1ull << 20;  // bytes per chunk, at least
constexpr unsigned kMaxThreads = 4;         // per call, the caller included

// One call's chunks: the caller waits until `remaining` reaches zero. A
// helper decrements it and notifies with `mu` held, so the caller (whose
// stack holds the job) cannot return before the helper is done with it.
struct Job {
    std::mutex mu;
    std::condition_variable done;
    unsigned remaining = 0;
};

struct Task {
    Job* job;
    uint8_t* dst;
    const uint8_t* src;
    uint64_t bytes;
};

class CopyPool {
public:
    // Never destroyed: its helpers wait for work until the process ends.
    static CopyPool& Get() {
        static CopyPool* pool = new CopyPool();
        return *pool;
    }

    unsigned threads() const { return static_cast<unsigned>(helpers_.size()) + 1; }

    void Copy(uint8_t* dst, const uint8_t* src, uint64_t bytes) {
        const uint64_t parts = std::min<uint64_t>(threads(), bytes / kMinChunk);
        // A forked child has none of the helpers (and the pool's lock may
        // have been held at the fork): it copies alone.
        if (parts < 2 || getpid() != pid_) {
            std::memcpy(dst, src, bytes);
            return;
        }
        // Chunk boundaries on 4 KiB multiples; the caller takes the first.
        const uint64_t chunk = ((bytes / parts) + 4095) & ~uint64_t{4095};
        Job job;
        std::vector<Task> tasks;
        for (uint64_t off = chunk; off < bytes; off += chunk)
            tasks.push_back(Task{&job, dst + off, src + off, std::min(chunk, bytes - off)});
        job.remaining = static_cast<unsigned>(tasks.size());
        {
            std::lock_guard<std::mutex> lock(mu_);
            for (const Task& t : tasks) queue_.push_back(t);
        }
        work_.notify_all();
        std::memcpy(dst, src, std::min(chunk, bytes));
        // Then take queued chunks (this call's or another's) until none is
        // left, so a call never waits on helpers busy elsewhere while work
        // it could do is queued.
        for (Task t; Pop(&t);) Run(t);
        std::unique_lock<std::mutex> lock(job.mu);
        job.done.wait(lock, [&] { return job.remaining == 0; });
    }

private:
    CopyPool() {
        cpu_set_t set;
        CPU_ZERO(&set);
        unsigned cpus = std::max(1u, std::thread::hardware_concurrency());
        if (sched_getaffinity(0, sizeof(set), &set) == 0) cpus = std::max(1, CPU_COUNT(&set));
        const unsigned helpers = std::min(kMaxThreads, cpus) - 1;
        for (unsigned i = 0; i < helpers; ++i) helpers_.emplace_back([this] { Help(); });
        for (std::thread& t : helpers_) t.detach();
    }

    bool Pop(Task* t) {
        std::lock_guard<std::mutex> lock(mu_);
        if (queue_.empty()) return false;
        *t = queue_.front();
        queue_.pop_front();
        return true;
    }

    static void Run(const Task& t) {
        std::memcpy(t.dst, t.src, t.bytes);
        std::lock_guard<std::mutex> lock(t.job->mu);
        if (--t.job->remaining == 0) t.job->done.notify_one();
    }

    void Help() {
        for (;;) {
            Task t;
            {
                std::unique_lock<std::mutex> lock(mu_);
                work_.wait(lock, [&] { return !queue_.empty(); });
                t = queue_.front();
                queue_.pop_front();
            }
            Run(t);
        }
    }

    const pid_t pid_ = getpid();
    std::mutex mu_;
    std::condition_variable work_;
    std::deque<Task> queue_;
    std::vector<std::thread> helpers_;
};

}  // namespace

void HostCopy(void* dst, const void* src, uint64_t bytes) {
    if (bytes < 2 * kMinChunk) {
        std::memcpy(dst, src, bytes);
        return;
    }
    CopyPool::Get().Copy(static_cast<uint8_t*>(dst), static_cast<const uint8_t*>(src), bytes);
}

unsigned HostCopyThreads() { return CopyPool::Get().threads(); }

}  // namespace vqm::ss2
