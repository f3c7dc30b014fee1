// Persistent worker threads for large host memory copies (pageable caller
// buffers <-> pinned staging), so a chunk copy uses several cores' memory
// bandwidth.  Plays the role of the reference's copy lanes
// (candidates/ordered-streaming/streaming_team.h:11-60, cpu_copy.h:14-39),
// without CPU pinning.
//
// Stores are non-temporal (AVX2 streaming stores where the CPU has them): the
// destinations -- pinned staging that a DMA engine reads next, the caller's
// output vector -- are not read back by this thread, so a cached store's
// read-for-ownership would only add a third of the bytes to host DRAM traffic,
// which is what bounds these copies.
//
// A product hands the pool one job per chunk (both operands' slices at once).
// post() starts a job and returns, so the caller can queue the previous
// chunk's transfer and kernels (HIP calls of 10-50 us each) while the workers
// copy; wait() then takes what is left and returns once every byte is copied.
// A job is cut into 256 KB pieces that the threads take from a shared counter.
// While a product is in flight (set_active(true)) an idle worker spins on the
// job counter for up to 300 us, so a hand-off between closely spaced chunks
// costs about a microsecond instead of a condition-variable wake-up; after
// that, or when no product is in flight, it sleeps (several lanes' workers
// would otherwise burn the cores that plan construction needs).
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <immintrin.h>
#include <mutex>
#include <thread>
#include <vector>

namespace pm {

__attribute__((target("avx2"))) inline void copy_streaming_avx2(char* out, const char* in, size_t bytes) {
    const size_t head = std::min(bytes, size_t((32 - (reinterpret_cast<uintptr_t>(out) & 31)) & 31));
    std::memcpy(out, in, head);
    out += head, in += head, bytes -= head;
    size_t i = 0;
    for (; i + 128 <= bytes; i += 128) {
        const __m256i v0 = _mm256_loadu_si256(reinterpret_cast<const __m256i*>(in + i));
        const __m256i v1 = _mm256_loadu_si256(reinterpret_cast<const __m256i*>(in + i + 32));
        const __m256i v2 = _mm256_loadu_si256(reinterpret_cast<const __m256i*>(in + i + 64));
        const __m256i v3 = _mm256_loadu_si256(reinterpret_cast<const __m256i*>(in + i + 96));
        _mm256_stream_si256(reinterpret_cast<__m256i*>(out + i), v0);
        _mm256_stream_si256(reinterpret_cast<__m256i*>(out + i + 32), v1);
        _mm256_stream_si256(reinterpret_cast<__m256i*>(out + i + 64), v2);
        _mm256_stream_si256(reinterpret_cast<__m256i*>(out + i + 96), v3);
    }
    std::memcpy(out + i, in + i, bytes - i);
}

// Copy with non-temporal stores, ordered (sfence) before the caller publishes
// completion to another thread or to a DMA engine.
inline void copy_streaming(void* dst, const void* src, size_t bytes) {
    static const bool avx2 = __builtin_cpu_supports("avx2");
    if (avx2)
        copy_streaming_avx2(static_cast<char*>(dst), static_cast<const char*>(src), bytes);
    else
        std::memcpy(dst, src, bytes);
    _mm_sfence();
}

class CopyPool {
public:
    struct Segment {
        void* dst;
        const void* src;
        size_t bytes;
    };

    explicit CopyPool(unsigned threads) : parts_(threads < 1 ? 1 : threads) {
        for (unsigned t = 1; t < parts_; ++t) workers_.emplace_back([this] { loop(); });
    }

    ~CopyPool() {
        wait();
        {
            std::lock_guard<std::mutex> lock(mutex_);
            stop_.store(true, std::memory_order_release);
            generation_.fetch_add(1, std::memory_order_acq_rel);
        }
        wake_.notify_all();
        for (auto& w : workers_) w.join();
    }

    CopyPool(const CopyPool&) = delete;
    CopyPool& operator=(const CopyPool&) = delete;

    unsigned threads() const { return parts_; }

    // A product starts (true: wake sleeping workers to spin for the first
    // chunk) or ends (false: idle workers sleep).  Not reentrant with copy().
    void set_active(bool on) {
        active_.store(on, std::memory_order_seq_cst);
        if (!on) return;
        pokes_.fetch_add(1, std::memory_order_seq_cst);
        if (sleepers_.load(std::memory_order_seq_cst) != 0) {
            std::lock_guard<std::mutex> lock(mutex_);
            wake_.notify_all();
        }
    }

    void copy(void* dst, const void* src, size_t bytes) {
        const Segment s{dst, src, bytes};
        copy(&s, 1);
    }

    // Copies every segment and returns when all are done.
    void copy(const Segment* segs, size_t count) {
        post(segs, count);
        wait();
    }

    // Starts copying the segments (the array itself may be reused at once) and
    // returns; jobs under 1 MB, or a pool of one thread, are copied here and
    // now.  One job at a time: wait() before the next post().
    void post(const Segment* segs, size_t count) {
        wait();
        size_t total = 0;
        for (size_t i = 0; i < count; ++i) total += segs[i].bytes;
        if (total < (size_t(1) << 20) || parts_ == 1) {
            for (size_t i = 0; i < count; ++i) std::memcpy(segs[i].dst, segs[i].src, segs[i].bytes);
            return;
        }
        pieces_.clear();
        for (size_t i = 0; i < count; ++i)
            for (size_t off = 0; off < segs[i].bytes; off += kPiece)
                pieces_.push_back({static_cast<char*>(segs[i].dst) + off, static_cast<const char*>(segs[i].src) + off,
                                   std::min(kPiece, segs[i].bytes - off)});
        next_.store(0, std::memory_order_relaxed);
        pending_.store(parts_ - 1, std::memory_order_relaxed);
        posted_ = true;
        // Publish, then wake sleepers.  A worker registers as a sleeper under
        // the mutex before it re-checks the counter, so either it sees this
        // job or this thread sees it registered (both sequentially consistent).
        generation_.fetch_add(1, std::memory_order_seq_cst);
        if (sleepers_.load(std::memory_order_seq_cst) != 0) {
            std::lock_guard<std::mutex> lock(mutex_);
            wake_.notify_all();
        }
    }

    // Takes pieces of the posted job until none are left, then waits until
    // every worker has finished its own and left the job (so the next post()
    // may rewrite the piece list).  Returns at once with no job posted.
    void wait() {
        if (!posted_) return;
        take();
        while (pending_.load(std::memory_order_acquire) != 0) pause();
        posted_ = false;
    }

private:
    static void pause() { __builtin_ia32_pause(); }

    // Copies pieces of the current job until the counter passes the last;
    // each piece's streaming stores are fenced before the thread moves on.
    void take() {
        const size_t count = pieces_.size();
        for (size_t i = next_.fetch_add(1, std::memory_order_relaxed); i < count;
             i = next_.fetch_add(1, std::memory_order_relaxed))
            copy_streaming(pieces_[i].dst, pieces_[i].src, pieces_[i].bytes);
    }

    void loop() {
        // the counter's initial value, not a fresh load: a job published before
        // this thread got here must still be taken
        size_t seen = 0;
        for (;;) {
            auto spin_until = std::chrono::steady_clock::now() + std::chrono::microseconds(300);
            unsigned polls = 0;
            while (generation_.load(std::memory_order_acquire) == seen && !stop_.load(std::memory_order_acquire)) {
                if (active_.load(std::memory_order_acquire) &&
                    ((++polls & 255) || std::chrono::steady_clock::now() < spin_until)) {
                    pause();
                    continue;
                }
                const size_t poke = pokes_.load(std::memory_order_seq_cst);
                std::unique_lock<std::mutex> lock(mutex_);
                sleepers_.fetch_add(1, std::memory_order_seq_cst);
                wake_.wait(lock, [&] {
                    return generation_.load(std::memory_order_seq_cst) != seen ||
                           stop_.load(std::memory_order_acquire) || pokes_.load(std::memory_order_seq_cst) != poke;
                });
                sleepers_.fetch_sub(1, std::memory_order_seq_cst);
                spin_until = std::chrono::steady_clock::now() + std::chrono::microseconds(300);
            }
            if (stop_.load(std::memory_order_acquire)) return;
            seen = generation_.load(std::memory_order_acquire);
            take();
            pending_.fetch_sub(1, std::memory_order_acq_rel);
        }
    }

    const unsigned parts_;
    std::vector<std::thread> workers_;
    std::mutex mutex_;
    std::condition_variable wake_;
    std::atomic<size_t> generation_{0};
    std::atomic<bool> active_{false};
    std::atomic<bool> stop_{false};
    std::atomic<unsigned> sleepers_{0};
    std::atomic<size_t> pokes_{0};
    static constexpr size_t kPiece = size_t(256) << 10;
    std::vector<Segment> pieces_;
    std::atomic<size_t> next_{0};
    std::atomic<unsigned> pending_{0};
    bool posted_ = false;  // the calling thread's own state
};

}  // namespace pm
