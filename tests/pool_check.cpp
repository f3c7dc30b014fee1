// pm::CopyPool under stress: jobs of random sizes (one or two segments, also
// below the pool's 1 MB threshold), products starting and ending at random
// (set_active), and pauses long enough for the workers to fall asleep between
// jobs, so every hand-off path runs -- spinning, sleeping, woken by a job,
// woken by a product start.  Half the jobs are posted and waited for later
// (post/wait), with the segment array overwritten in between.  Every copy is
// checked.  Prints "ok <jobs>".
#include <chrono>
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "copy_pool.hpp"

int main(int argc, char** argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 1500;
    std::mt19937 rng(7);
    int jobs = 0;
    for (unsigned threads : {1u, 2u, 4u, 8u}) {
        pm::CopyPool pool(threads);
        std::vector<uint8_t> src(size_t(8) << 20), dst(src.size()), dst2(size_t(1) << 20);
        for (auto& x : src) x = uint8_t(rng());
        for (int it = 0; it < iters; ++it, ++jobs) {
            pool.set_active(rng() % 3 != 0);
            const size_t n = rng() % 8 == 0 ? rng() % (size_t(1) << 20) : (size_t(1) << 20) + rng() % (size_t(7) << 20);
            const size_t n2 = rng() % 2 ? rng() % dst2.size() : 0;
            for (size_t i = 0; i < 64; ++i) src[(i * 131071) % src.size()] = uint8_t(it + i);  // fresh contents
            pm::CopyPool::Segment segs[2] = {{dst.data(), src.data(), n}, {dst2.data(), src.data() + 3, n2}};
            if (rng() % 2) {
                pool.copy(segs, n2 ? 2 : 1);
            } else {
                // posted: the segment array is reused at once, the caller works
                // meanwhile, and wait() (sometimes called twice) finishes the job
                pool.post(segs, n2 ? 2 : 1);
                segs[0] = segs[1] = {nullptr, nullptr, 0};
                if (rng() % 3 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 200));
                pool.wait();
                if (rng() % 5 == 0) pool.wait();
            }
            if (std::memcmp(dst.data(), src.data(), n) || std::memcmp(dst2.data(), src.data() + 3, n2)) {
                std::printf("MISMATCH threads=%u job=%d bytes=%zu+%zu\n", threads, it, n, n2);
                return 1;
            }
            if (rng() % 40 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 1500));
            if (rng() % 4 == 0) pool.set_active(false);
        }
    }
    std::printf("ok %d\n", jobs);
    return 0;
}
