// Process-wide cache of pinned host blocks (hipHostMalloc).
//
// Fresh pinned memory is populated on its first CPU write, at about 70 us per
// MB on the MI355X host (a 1.2 GB staging set costs some 85 ms on the first
// product that uses it, and prefaulting from several threads costs the same),
// while later writes run at memory speed.  Plans and lanes therefore take
// their staging from here and give it back when they are destroyed or grow:
// the next plan of a similar size reuses populated memory.  A request takes
// the smallest cached block that holds it, if that block is at most twice its
// size; otherwise a new block is allocated (if pinning fails, every cached
// block is freed and the allocation tried once more).  Cached bytes are capped
// (a quarter of physical memory, at most 32 GiB); beyond that the largest
// cached blocks are freed.  Blocks still cached at exit are left to the OS.
#pragma once

#include <hip/hip_runtime.h>
#include <unistd.h>

#include <algorithm>
#include <cstddef>
#include <map>
#include <mutex>
#include <new>

namespace pm {

class PinnedCache {
public:
    static PinnedCache& instance() {
        static PinnedCache* cache = new PinnedCache();  // never destroyed: the HIP runtime may be gone at exit
        return *cache;
    }

    // A pinned block of at least `bytes`; `*got` receives its size.
    // Throws std::bad_alloc when the runtime cannot pin the memory.
    void* take(size_t bytes, size_t* got) {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            auto it = blocks_.lower_bound(bytes);
            if (it != blocks_.end() && it->first <= 2 * bytes) {
                void* p = it->second;
                *got = it->first;
                cached_ -= it->first;
                blocks_.erase(it);
                return p;
            }
        }
        void* p = allocate(bytes);
        if (!p) {  // pinned memory exhausted: give back what the cache holds and try once more
            release_all();
            p = allocate(bytes);
            if (!p) throw std::bad_alloc();
        }
        *got = bytes;
        return p;
    }

    void give(void* p, size_t bytes) {
        if (!p) return;
        std::lock_guard<std::mutex> lock(mutex_);
        blocks_.emplace(bytes, p);
        cached_ += bytes;
        while (cached_ > cap_ && !blocks_.empty()) {
            auto largest = std::prev(blocks_.end());
            cached_ -= largest->first;
            (void)hipHostFree(largest->second);
            blocks_.erase(largest);
        }
    }

private:
    static void* allocate(size_t bytes) {
        void* p = nullptr;
        if (hipHostMalloc(&p, bytes, hipHostMallocDefault) == hipSuccess && p) return p;
        (void)hipGetLastError();  // clear the failure so later calls do not report it
        return nullptr;
    }

    void release_all() {
        std::lock_guard<std::mutex> lock(mutex_);
        for (auto& [size, block] : blocks_) (void)hipHostFree(block);
        blocks_.clear();
        cached_ = 0;
    }

    PinnedCache() {
        const long pages = sysconf(_SC_PHYS_PAGES), page = sysconf(_SC_PAGE_SIZE);
        const size_t quarter = pages > 0 && page > 0 ? size_t(pages) * size_t(page) / 4 : size_t(4) << 30;
        cap_ = std::min(quarter, size_t(32) << 30);
    }

    std::mutex mutex_;
    std::multimap<size_t, void*> blocks_;
    size_t cached_ = 0, cap_ = 0;
};

}  // namespace pm
