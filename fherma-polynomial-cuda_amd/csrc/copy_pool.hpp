// Persistent worker threads for large host memory copies (pageable caller
// buffers <-> pinned staging), so a chunk copy uses several cores' memory
// bandwidth.  Plays the role of the reference's copy lanes
// (candidates/ordered-streaming/streaming_team.h:11-60, cpu_copy.h:14-39),
// without CPU pinning or non-temporal stores.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace pm {

class CopyPool {
public:
    explicit CopyPool(unsigned threads) : parts_(threads < 1 ? 1 : threads) {
        for (unsigned t = 1; t < parts_; ++t) workers_.emplace_back([this, t] { loop(t); });
    }

    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            stop_ = true;
            ++generation_;
        }
        wake_.notify_all();
        for (auto& w : workers_) w.join();
    }

    CopyPool(const CopyPool&) = delete;
    CopyPool& operator=(const CopyPool&) = delete;

    // Copies bytes from src to dst; the calling thread takes one share.
    // Returns when every share is done.  Not reentrant: one job at a time.
    void copy(void* dst, const void* src, size_t bytes) {
        if (bytes < (size_t(1) << 20) || parts_ == 1) {
            std::memcpy(dst, src, bytes);
            return;
        }
        {
            std::lock_guard<std::mutex> lock(mutex_);
            dst_ = static_cast<char*>(dst);
            src_ = static_cast<const char*>(src);
            bytes_ = bytes;
            pending_.store(parts_ - 1, std::memory_order_relaxed);
            ++generation_;
        }
        wake_.notify_all();
        share(0);
        std::unique_lock<std::mutex> lock(mutex_);
        done_.wait(lock, [this] { return pending_.load(std::memory_order_acquire) == 0; });
    }

    unsigned threads() const { return parts_; }

private:
    void share(unsigned t) {
        // 4 KB-aligned slices so neighbouring shares never split a page
        const size_t pages = (bytes_ + 4095) / 4096;
        const size_t first = pages * t / parts_ * 4096, last = std::min(bytes_, pages * (t + 1) / parts_ * 4096);
        if (last > first) std::memcpy(dst_ + first, src_ + first, last - first);
    }

    void loop(unsigned t) {
        size_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lock(mutex_);
                wake_.wait(lock, [&] { return generation_ != seen; });
                seen = generation_;
                if (stop_) return;
            }
            share(t);
            if (pending_.fetch_sub(1, std::memory_order_acq_rel) == 1) {
                std::lock_guard<std::mutex> lock(mutex_);
                done_.notify_one();
            }
        }
    }

    const unsigned parts_;
    std::vector<std::thread> workers_;
    std::mutex mutex_;
    std::condition_variable wake_, done_;
    size_t generation_ = 0;
    bool stop_ = false;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0;
    std::atomic<unsigned> pending_{0};
};

}  // namespace pm
