// Coefficient chunks of the host-vector pipeline (no HIP dependence, so the
// CPU tests include it directly).  The pipeline moves operands in and results
// out in up to kMaxChunks chunks of whole 256-coefficient tiles, so that host
// copies, transfers and the per-coefficient stages of neighbouring chunks overlap.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>

namespace pm {

constexpr int kMaxChunks = 12;

struct Chunking {
    uint32_t count = 1;   // chunks
    uint32_t ntiles = 1;  // 256-coefficient tiles in all
    // Chunk ch holds tiles [start[ch], start[ch + 1]); every chunk is non-empty.
    uint32_t start[kMaxChunks + 1] = {};
    // Tiles [t0, t0 + tc) of chunk ch; tc >= 1 for every ch < count.
    void tiles(uint32_t ch, uint32_t& t0, uint32_t& tc) const {
        t0 = start[ch];
        tc = start[ch + 1] - start[ch];
    }
    // Coefficient rows [r0, r1) of chunk ch (the last tile may be partly past n).
    void rows(uint32_t ch, size_t n, size_t& r0, size_t& r1) const {
        r0 = std::min<size_t>(size_t(start[ch]) * 256, n);
        r1 = std::min<size_t>(size_t(start[ch + 1]) * 256, n);
    }
};

// About one chunk per 6 MB of operand, up to 8 of `per` tiles each: each chunk
// costs some tens of microseconds of copy-engine and event hand-offs, which
// small products feel, while large ones want the overlap of many chunks.
// With 4 or more such chunks the two at each end are graded to a quarter and a
// half of the others: the first chunk's host copy comes before any transfer
// starts, and after the last transfer the last chunk's conversion, the first
// output chunk's CRT GEMM and, at the very end, the last chunk's copy into the
// caller's vector all run with the link idle -- the smaller those chunks, the
// less of that is exposed.  Every chunk starts below the last tile.
inline Chunking make_chunking(uint32_t n_pad, uint32_t n, uint32_t limbs) {
    const uint32_t ntiles = std::max(1u, n_pad / 256);
    const size_t bytes = size_t(n) * limbs * sizeof(uint32_t);
    const size_t unit = size_t(6) << 20;
    const uint32_t limit = uint32_t(std::max<size_t>(1, std::min<size_t>(8, (bytes + unit - 1) / unit)));
    const uint32_t want = std::max(1u, std::min(ntiles, limit));
    const uint32_t per = (ntiles + want - 1) / want;
    Chunking c;
    c.ntiles = ntiles;
    uint32_t sizes[kMaxChunks];
    uint32_t count = 0;
    if (want >= 4 && per >= 8) {
        const uint32_t quarter = per / 4, half = per / 2;
        uint32_t middle = ntiles - 2 * (quarter + half);  // ntiles > 3 per: positive
        sizes[count++] = quarter;
        sizes[count++] = half;
        while (middle) {
            const uint32_t take = std::min(per, middle);
            sizes[count++] = take;
            middle -= take;
        }
        sizes[count++] = half;
        sizes[count++] = quarter;
    } else {
        for (uint32_t t = 0; t < ntiles; t += per) sizes[count++] = std::min(per, ntiles - t);
    }
    c.count = count;
    c.start[0] = 0;
    for (uint32_t i = 0; i < count; ++i) c.start[i + 1] = c.start[i] + sizes[i];
    return c;
}

}  // namespace pm
