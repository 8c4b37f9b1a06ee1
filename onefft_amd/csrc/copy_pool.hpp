// A fork-join pool for large host-memory copies. Plain C++ with no HIP: the host staging path
// (staging.hip) uses it, and a CPU test (tests/test_copy_pool.py) compiles it on its own.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace onefft {

// The ring's CPU copies, split over a few threads: one thread copies pageable memory at about
// half the rate the DMA engine moves pinned memory over the host link, so a single-threaded ring
// is bound by its memcpy. A process-wide fork-join pool of ONEFFT_COPY_THREADS helper threads
// (default 4, 0 for none) copies 1 MiB pieces of a chunk together with the calling thread; one
// job runs at a time, and a call from another thread meanwhile copies alone instead of waiting.
// Each job lives on its caller's stack and is released only once no helper refers to it.
class CopyPool {
public:
    static CopyPool &get() {
        static CopyPool pool(threads_from_env());
        return pool;
    }

    // a pool of `helpers` threads (0..16) besides the callers
    explicit CopyPool(int helpers) {
        if (helpers < 0) helpers = 0;
        if (helpers > 16) helpers = 16;
        for (int i = 0; i < helpers; i++) helpers_.emplace_back([this] { helper(); });
    }

    void copy(void *dst, const void *src, size_t bytes) {
        if (helpers_.empty() || bytes <= kPiece) {
            std::memcpy(dst, src, bytes);
            return;
        }
        // the helpers serve one job at a time; a call that finds them taken copies alone rather
        // than wait for another caller's copy
        std::unique_lock<std::mutex> turn(turn_, std::try_to_lock);
        if (!turn.owns_lock()) {
            std::memcpy(dst, src, bytes);
            return;
        }
        Job job{static_cast<char *>(dst), static_cast<const char *>(src), bytes, (bytes + kPiece - 1) / kPiece};
        {
            std::lock_guard<std::mutex> lk(m_);
            job_ = &job;
            ++gen_;
        }
        wake_.notify_all();
        const size_t mine = run(job);
        std::unique_lock<std::mutex> lk(m_);
        job.finished += mine;
        job_ = nullptr;  // late helpers find no job
        done_.wait(lk, [&] { return job.finished == job.pieces && job.active == 0; });
    }

    CopyPool(const CopyPool &) = delete;
    CopyPool &operator=(const CopyPool &) = delete;

    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
            ++gen_;
        }
        wake_.notify_all();
        for (auto &t : helpers_) t.join();
    }

private:
    static constexpr size_t kPiece = size_t(1) << 20;

    struct Job {
        char *dst;
        const char *src;
        size_t bytes, pieces;
        std::atomic<size_t> next{0};
        size_t finished = 0;  // pieces copied (under m_)
        int active = 0;       // helpers working on this job (under m_)
    };

    static int threads_from_env() {
        const char *env = std::getenv("ONEFFT_COPY_THREADS");
        return env ? std::atoi(env) : 4;
    }

    static size_t run(Job &job) {
        size_t count = 0;
        for (;;) {
            const size_t i = job.next.fetch_add(1, std::memory_order_relaxed);
            if (i >= job.pieces) return count;
            const size_t off = i * kPiece, len = job.bytes - off < kPiece ? job.bytes - off : kPiece;
            std::memcpy(job.dst + off, job.src + off, len);
            count++;
        }
    }

    void helper() {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            wake_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
            if (stop_) return;
            Job *job = job_;
            if (!job) continue;
            job->active++;
            lk.unlock();
            const size_t mine = run(*job);
            lk.lock();
            job->finished += mine;
            job->active--;
            done_.notify_all();
        }
    }

    std::vector<std::thread> helpers_;
    std::mutex turn_, m_;
    std::condition_variable wake_, done_;
    Job *job_ = nullptr;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

}  // namespace onefft
