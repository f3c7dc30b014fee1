// pm::BasicDeviceArena on the CPU, with a fake allocator (malloc under a
// simulated device capacity that can refuse): idle chunks too small for a
// request must not block it, admission waits for a release, requests beyond
// the budget or the device fail, freed neighbours coalesce, and under threads
// no two live blocks overlap nor does their sum pass the budget.  Prints
// "ok" and exits 0 when every check holds.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "arena.hpp"

namespace {

struct FakeDevice {
    static inline std::mutex mutex;
    static inline size_t capacity = 0, used = 0;
    static inline std::map<void*, size_t> blocks;
    static bool allocate(void** p, size_t bytes) {
        std::lock_guard<std::mutex> lock(mutex);
        if (used + bytes > capacity) return false;
        *p = std::malloc(bytes);
        if (!*p) return false;
        used += bytes;
        blocks[*p] = bytes;
        return true;
    }
    static void release(void* p) {
        std::lock_guard<std::mutex> lock(mutex);
        used -= blocks.at(p);
        blocks.erase(p);
        std::free(p);
    }
};

using Arena = pm::BasicDeviceArena<FakeDevice>;
int failures = 0;
#define CHECK(cond)                                                     \
    do {                                                                \
        if (!(cond)) {                                                  \
            std::printf("FAIL line %d: %s\n", __LINE__, #cond);         \
            ++failures;                                                 \
        }                                                               \
    } while (0)

bool throws_bad_alloc(Arena& a, size_t bytes) {
    try {
        a.acquire(bytes);
    } catch (const std::bad_alloc&) {
        return true;
    }
    return false;
}

}  // namespace

int main() {
    const size_t K = 1 << 16;  // the test's chunk unit
    FakeDevice::capacity = 1 << 30;
    {  // idle chunks too small for a request hold the whole budget: they are given back
        Arena a(4 * K, K);
        std::vector<char*> p;
        for (int i = 0; i < 4; ++i) p.push_back(a.acquire(K));
        CHECK(a.allocated() == 4 * K);
        for (char* x : p) a.release(x);
        char* big = a.acquire(2 * K);
        CHECK(big != nullptr);
        CHECK(a.allocated() <= 4 * K);
        a.release(big);
    }
    {  // beyond the budget: fails at once
        Arena a(4 * K, K);
        CHECK(throws_bad_alloc(a, 4 * K + 256));
    }
    {  // admission: a request waits until a release makes room, then proceeds
        Arena a(3 * K, 3 * K);
        char* first = a.acquire(2 * K);
        std::atomic<bool> released{false}, got{false};
        std::thread t([&] {
            char* second = a.acquire(2 * K);
            got = true;
            CHECK(released.load());
            a.release(second);
        });
        std::this_thread::sleep_for(std::chrono::milliseconds(50));
        CHECK(!got.load());
        released = true;
        a.release(first);
        t.join();
        CHECK(got.load());
    }
    {  // coalescing: freed neighbours merge, so the span is reused without growing
        Arena a(2 * K, K);
        char* x = a.acquire(K / 4);
        char* y = a.acquire(K / 4);
        char* z = a.acquire(K / 4);
        CHECK(y == x + K / 4 && z == y + K / 4);
        a.release(y);
        a.release(x);
        char* w = a.acquire(K / 2);
        CHECK(w == x);
        CHECK(a.allocated() == K);
        a.release(w);
        a.release(z);
    }
    {  // the device refuses more than it has: nothing in use -> bad_alloc; otherwise idle chunks go back
        FakeDevice::capacity = 2 * K;
        Arena a(8 * K, K);
        CHECK(throws_bad_alloc(a, 3 * K));
        char* x = a.acquire(K);
        a.release(x);  // one idle chunk of K
        char* y = a.acquire(2 * K);  // needs the idle chunk given back to fit the device
        CHECK(y != nullptr);
        a.release(y);
        FakeDevice::capacity = 1 << 30;
    }
    {  // threads: live blocks never overlap and never pass the budget
        const size_t budget = 16 * K;
        Arena a(budget, 2 * K);
        std::mutex m;
        std::map<char*, size_t> live;
        size_t live_bytes = 0;
        std::vector<std::thread> threads;
        for (int t = 0; t < 8; ++t)
            threads.emplace_back([&, t] {
                std::mt19937 rng(100 + t);
                for (int i = 0; i < 400; ++i) {
                    const size_t bytes = pm::aligned_bytes(256 + rng() % (5 * K));
                    char* p = a.acquire(bytes);
                    {
                        std::lock_guard<std::mutex> lock(m);
                        auto next = live.lower_bound(p);
                        if (next != live.end()) CHECK(p + bytes <= next->first);
                        if (next != live.begin()) {
                            auto prev = std::prev(next);
                            CHECK(prev->first + prev->second <= p);
                        }
                        live[p] = bytes;
                        live_bytes += bytes;
                        CHECK(live_bytes <= budget);
                    }
                    if (rng() % 8 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 300));
                    {
                        std::lock_guard<std::mutex> lock(m);
                        live.erase(p);
                        live_bytes -= bytes;
                    }
                    a.release(p);
                }
            });
        for (auto& th : threads) th.join();
        CHECK(a.allocated() <= budget);
    }
    CHECK(FakeDevice::used == 0);  // every chunk went back with its arena
    if (failures) return 1;
    std::printf("ok\n");
    return 0;
}
