// The host pipeline's coefficient chunks (csrc/chunks.hpp) for every
// power-of-two N up to 2^24 and every limb count up to 8192 (beyond the
// 16383-prime cap): chunk k's tile range lies inside [0, ntiles), the chunks
// cover every tile exactly once in order, no chunk is empty, at most
// kMaxChunks are used, and the row ranges cover [0, N).  Prints "ok <cases>".
#include <cstdio>
#include <cstdlib>

#include "chunks.hpp"

int main() {
    unsigned long cases = 0;
    for (unsigned log_n = 0; log_n <= 24; ++log_n) {
        const uint32_t n = 1u << log_n, n_pad = (n + 255) / 256 * 256;
        for (uint32_t limbs = 1; limbs <= 8192; ++limbs) {
            const pm::Chunking c = pm::make_chunking(n_pad, n, limbs);
            if (c.count < 1 || c.count > uint32_t(pm::kMaxChunks) || c.ntiles != n_pad / 256) {
                std::printf("bad count: N=%u L=%u count=%u ntiles=%u\n", n, limbs, c.count, c.ntiles);
                return 1;
            }
            uint32_t next_tile = 0;
            size_t next_row = 0;
            for (uint32_t ch = 0; ch < c.count; ++ch) {
                uint32_t t0, tc;
                c.tiles(ch, t0, tc);
                size_t r0, r1;
                c.rows(ch, n, r0, r1);
                if (t0 != next_tile || tc == 0 || t0 + tc > c.ntiles || r0 != next_row || r1 <= r0 ||
                    r0 != size_t(t0) * 256 || r1 != std::min<size_t>(size_t(t0 + tc) * 256, n)) {
                    std::printf("bad chunk: N=%u L=%u chunk %u/%u tiles [%u,+%u) of %u rows [%zu,%zu)\n", n, limbs,
                                ch, c.count, t0, tc, c.ntiles, r0, r1);
                    return 1;
                }
                next_tile = t0 + tc;
                next_row = r1;
            }
            if (next_tile != c.ntiles || next_row != n) {
                std::printf("chunks do not cover: N=%u L=%u\n", n, limbs);
                return 1;
            }
            ++cases;
        }
    }
    std::printf("ok %lu\n", cases);
    return 0;
}
