// Coefficient chunks of the host-vector pipeline (no HIP dependence, so the
// CPU tests include it directly).  The pipeline moves operands in and results
// out in up to kMaxChunks chunks of whole 256-coefficient tiles, so that host
// copies, transfers and the per-coefficient stages of neighbouring chunks overlap.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>

namespace pm {

constexpr int kMaxChunks = 8;

struct Chunking {
    uint32_t count = 1, per = 1;  // chunks, tiles per chunk (every chunk starts below the last tile)
    uint32_t ntiles = 1;          // 256-coefficient tiles in all
    // Tiles [t0, t0 + tc) of chunk ch; tc >= 1 for every ch < count.
    void tiles(uint32_t ch, uint32_t& t0, uint32_t& tc) const {
        t0 = ch * per;
        tc = t0 < ntiles ? std::min(per, ntiles - t0) : 0;
    }
    // Coefficient rows [r0, r1) of chunk ch (the last tile may be partly past n).
    void rows(uint32_t ch, size_t n, size_t& r0, size_t& r1) const {
        r0 = std::min<size_t>(size_t(ch) * per * 256, n);
        r1 = std::min<size_t>(size_t(ch + 1) * per * 256, n);
    }
};

// About one chunk per 6 MB of operand, 1 .. kMaxChunks: each chunk costs some
// tens of microseconds of copy-engine and event hand-offs, which small products
// feel, while large ones want the overlap of many chunks.  The count is then
// trimmed to ceil(ntiles / per), so that no chunk starts past the last tile
// (8 wanted chunks of 8 tiles need 7 chunks of 2, not 8).
inline Chunking make_chunking(uint32_t n_pad, uint32_t n, uint32_t limbs) {
    const uint32_t ntiles = std::max(1u, n_pad / 256);
    const size_t bytes = size_t(n) * limbs * sizeof(uint32_t);
    const size_t unit = size_t(6) << 20;
    const uint32_t limit = uint32_t(std::max<size_t>(1, std::min<size_t>(kMaxChunks, (bytes + unit - 1) / unit)));
    const uint32_t want = std::max(1u, std::min(ntiles, limit));
    Chunking c;
    c.ntiles = ntiles;
    c.per = (ntiles + want - 1) / want;
    c.count = (ntiles + c.per - 1) / c.per;
    return c;
}

}  // namespace pm
