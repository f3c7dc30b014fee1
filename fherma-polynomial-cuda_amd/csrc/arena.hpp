// The device memory arena of a server (see BasicDeviceArena).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstddef>
#include <iterator>
#include <map>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

namespace pm {

// Device allocations are rounded to 256 bytes (GEMM tiles load 16-byte vectors;
// every buffer starts on a 256-byte boundary).
inline size_t aligned_bytes(size_t bytes) { return ((bytes ? bytes : 1) + 255) & ~size_t(255); }

// Device memory of a server: at most `budget` bytes, taken from the allocator
// (hipMalloc) in chunks as requests need them (a chunk holds at least
// `min_chunk` -- 4 GiB -- or the request, never past the budget) and handed
// out first-fit within the chunks, with coalescing.  acquire() blocks while no
// chunk has room and a new chunk would pass the budget: that is the admission
// rule.  Chunks with nothing in use are given back when the device refuses a
// new chunk, and when idle chunks too small for a request hold the budget it
// needs; a request that still cannot be placed while nothing of the arena's is
// in use fails with std::bad_alloc.  (Reserving the whole budget up front cost
// from a few ms to 0.7 s per server in hipMalloc, and up to 0.8 s more in
// hipFree, on the MI355X.)  The chunks are freed with the arena.
//
// Alloc provides `static bool allocate(void** p, size_t bytes)` (false: out of
// memory), `static void release(void* p)`; the default is the HIP runtime.
struct HipAllocator {
    static bool allocate(void** p, size_t bytes) {
        const hipError_t e = hipMalloc(p, bytes);
        if (e == hipSuccess) return true;
        (void)hipGetLastError();  // clear the error of a refused allocation
        if (e == hipErrorOutOfMemory) return false;
        throw std::runtime_error(std::string("arena hipMalloc: ") + hipGetErrorString(e));
    }
    static void release(void* p) { (void)hipFree(p); }
};

template <class Alloc = HipAllocator>
class BasicDeviceArena {
public:
    explicit BasicDeviceArena(size_t budget, size_t min_chunk = size_t(4) << 30)
        : budget_(budget & ~size_t(255)), min_chunk_(min_chunk) {}
    ~BasicDeviceArena() {
        for (Chunk& c : chunks_) Alloc::release(c.base);
    }
    BasicDeviceArena(const BasicDeviceArena&) = delete;
    BasicDeviceArena& operator=(const BasicDeviceArena&) = delete;

    size_t capacity() const { return budget_; }
    size_t allocated() const {
        std::lock_guard<std::mutex> lock(mutex_);
        return allocated_;
    }

    char* acquire(size_t bytes) {
        bytes = aligned_bytes(bytes);
        if (bytes > budget_) throw std::bad_alloc();
        std::unique_lock<std::mutex> lock(mutex_);
        for (;;) {
            if (char* p = carve(bytes)) return p;
            // Idle chunks too small for this request still count against the
            // budget: give them back before concluding it must wait (else a
            // request could wait for a release that never makes room).
            if (allocated_ + bytes > budget_) drop_idle_chunks();
            if (allocated_ + bytes <= budget_ && grow(bytes)) continue;
            if (used_.empty()) throw std::bad_alloc();  // nothing will be released to make room
            released_.wait(lock);
        }
    }

    void release(char* p) {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            auto it = used_.find(p);
            if (it == used_.end()) return;
            Chunk& c = chunks_[it->second.chunk];
            size_t off = size_t(p - c.base), size = it->second.bytes;
            used_.erase(it);
            --c.in_use;
            auto next = c.free.lower_bound(off);
            if (next != c.free.end() && next->first == off + size) {
                size += next->second;
                next = c.free.erase(next);
            }
            if (next != c.free.begin()) {
                auto prev = std::prev(next);
                if (prev->first + prev->second == off) {
                    prev->second += size;
                    size = 0;
                }
            }
            if (size) c.free[off] = size;
        }
        released_.notify_all();
    }

private:
    struct Chunk {
        char* base;
        size_t size;
        size_t in_use = 0;
        std::map<size_t, size_t> free;  // offset -> size
    };
    struct Use {
        size_t chunk, bytes;
    };

    char* carve(size_t bytes) {  // first fit; caller holds the mutex
        for (size_t ci = 0; ci < chunks_.size(); ++ci) {
            Chunk& c = chunks_[ci];
            for (auto it = c.free.begin(); it != c.free.end(); ++it) {
                if (it->second < bytes) continue;
                const size_t off = it->first, size = it->second;
                c.free.erase(it);
                if (size > bytes) c.free[off + bytes] = size - bytes;
                ++c.in_use;
                used_[c.base + off] = Use{ci, bytes};
                return c.base + off;
            }
        }
        return nullptr;
    }

    bool grow(size_t bytes) {  // caller holds the mutex
        for (int attempt = 0; attempt < 2; ++attempt) {
            const size_t size = std::min(budget_ - allocated_, std::max(bytes, min_chunk_));
            void* base = nullptr;
            if (Alloc::allocate(&base, size)) {
                chunks_.push_back(Chunk{static_cast<char*>(base), size});
                chunks_.back().free[0] = size;
                allocated_ += size;
                return true;
            }
            if (attempt == 0 && !drop_idle_chunks()) break;
        }
        return false;
    }

    bool drop_idle_chunks() {  // frees chunks with nothing in use; caller holds the mutex
        bool dropped = false;
        for (size_t ci = chunks_.size(); ci-- > 0;) {
            if (chunks_[ci].in_use) continue;
            Alloc::release(chunks_[ci].base);
            allocated_ -= chunks_[ci].size;
            chunks_.erase(chunks_.begin() + ci);
            dropped = true;
        }
        if (dropped)  // chunk indices moved: renumber the blocks in use
            for (auto& u : used_)
                for (size_t ci = 0; ci < chunks_.size(); ++ci)
                    if (u.first >= chunks_[ci].base && u.first < chunks_[ci].base + chunks_[ci].size) u.second.chunk = ci;
        return dropped;
    }

    size_t budget_, min_chunk_, allocated_ = 0;
    mutable std::mutex mutex_;
    std::condition_variable released_;
    std::vector<Chunk> chunks_;
    std::map<char*, Use> used_;
};

using DeviceArena = BasicDeviceArena<>;

}  // namespace pm
