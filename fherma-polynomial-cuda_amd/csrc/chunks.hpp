// Coefficient chunks of the host-vector pipeline (no HIP dependence, so the
// CPU tests include it directly).  The pipeline moves operands in and results
// out in up to kMaxChunks chunks of whole 256-coefficient tiles, so that host
// copies, transfers and the per-coefficient stages of neighbouring chunks overlap.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>

namespace pm {

constexpr int kMaxChunks = 16;
// Runs of consecutive tiles in one chunk: at most the 64 rows of a strip.
constexpr uint32_t kMaxRuns = 64;

// Tiles t0 + (k / w) * stride + k % w for k < count (count a multiple of w):
// count / w runs of w consecutive tiles, run g starting g * stride tiles after
// run 0.  A range of consecutive tiles is one run (w = count).
struct TileSet {
    uint32_t t0 = 0, count = 0, w = 1, stride = 0;
};

// The k-th tile of a set (constexpr, so device code calls it too).
constexpr uint32_t tile_at(const TileSet& s, uint32_t k) { return s.t0 + (k / s.w) * s.stride + k % s.w; }

// The last tile of a non-empty set.
constexpr uint32_t last_tile(const TileSet& s) { return tile_at(s, s.count - 1); }

constexpr TileSet tile_range(uint32_t t0, uint32_t count) { return TileSet{t0, count, count ? count : 1, 0}; }

struct Chunking {
    uint32_t count = 1;   // chunks
    uint32_t ntiles = 1;  // 256-coefficient tiles in all
    // Column strips: the N coefficients are the [rows][row_width] matrix the
    // transforms use (coefficient i = row * row_width + column), and a strip is
    // 256 consecutive columns over all rows, i.e. one tile per row.  Chunk ch
    // then holds a range of rows of one strip (all of them, or a piece); the forward
    // column stages of strips [fwd_begin[ch], fwd_end[ch]) can run once chunk ch
    // is in, and the inverse column stages of strips [inv_begin[ch],
    // inv_end[ch]) must run before chunk ch's output.  Without strips, chunks
    // are consecutive tile ranges and the column stages run whole.
    bool strips = false;
    TileSet set[kMaxChunks];
    // Chunk ch is rows [local[ch], local[ch + 1]) of the chunk-major staging
    // order (the operands' rows of chunk 0, then chunk 1, ...); local[count] = N.
    uint32_t local[kMaxChunks + 1] = {};
    uint32_t fwd_begin[kMaxChunks] = {}, fwd_end[kMaxChunks] = {};
    uint32_t inv_begin[kMaxChunks] = {}, inv_end[kMaxChunks] = {};

    uint32_t rows_of(uint32_t ch) const { return local[ch + 1] - local[ch]; }
    uint32_t runs(uint32_t ch) const { return set[ch].count / set[ch].w; }
    // Run g of chunk ch: coefficients [global0, global0 + len) of the operands,
    // which are rows [local0, local0 + len) of the chunk-major order.
    void run(uint32_t ch, uint32_t g, size_t n, size_t& global0, size_t& local0, size_t& len) const {
        const TileSet& s = set[ch];
        global0 = std::min<size_t>(size_t(tile_at(s, g * s.w)) * 256, n);
        const size_t end = std::min<size_t>(size_t(tile_at(s, g * s.w) + s.w) * 256, n);
        len = end - global0;
        local0 = local[ch] + size_t(g) * s.w * 256;
    }
};

// About one chunk per 6 MB of operand, up to 8 of `per` tiles each: each chunk
// costs some tens of microseconds of copy-engine and event hand-offs, which
// small products feel, while large ones want the overlap of many chunks.
//
// Column strips are used when the transform has a column stage held in
// registers (2 to 64 rows of a multiple of 256 columns, so at most 8 strips)
// and at least one chunk per strip is wanted: the forward column stages of a
// strip then run while later strips are still on the link, and the inverse
// column stages of a strip run while earlier strips' results go out.  With 4
// rows or more the end strips go in pieces of rows: the first and the last in
// three (a quarter, a quarter, a half; a half, a quarter, a quarter), the
// second and the second to last in halves, the rest whole (at 8 strips: 14
// chunks).  The first piece's host copy comes before any transfer starts, and
// after the last transfer the last piece's conversion, its strip's forward
// column stages and the row stage, and on the way out the first piece's
// inverse column stages and CRT GEMM, run with the link idle -- the smaller
// those pieces, the less of that is exposed.  The halves keep the compute
// stream from trailing the link: a chunk's conversion starts when its transfer
// ends, so a whole strip just before the last one (its conversion takes 0.68
// of its transfer time) would delay the last pieces' conversions; and on the
// way out a strip's inverse stages, CRT GEMM and reconstruction (0.83 of its
// D2H time) must fit under the previous chunk's D2H.  (Chunks of several
// whole strips, for fewer chunks, measured slower than tile ranges at
// N = 65536 with 55 limbs: 1.35-1.38 against 1.28-1.29 ms per product.)
//
// Without strips: consecutive tile ranges, and with 4 or more chunks the two at
// each end are graded to a quarter and a half of the others, for the same
// reason.  With 3 the last range is half the others: at N = 65536 with 55
// limbs, ranges of 104, 104 and 48 tiles measured 1.18-1.19 ms per product
// against 1.21 for 86, 86 and 84, while grading both ends, or more chunks,
// measured slower (each chunk still costs tens of microseconds of hand-offs).
// With 2, equal halves measured best (N = 32768, 55 limbs: 0.68 ms against
// 0.69-0.70 for 86 and 42 tiles).  Every chunk starts below the last tile.
inline Chunking make_chunking(uint32_t n_pad, uint32_t n, uint32_t limbs, uint32_t row_width = 1,
                              uint32_t rows = 1) {
    const uint32_t ntiles = std::max(1u, n_pad / 256);
    const size_t bytes = size_t(n) * limbs * sizeof(uint32_t);
    const size_t unit = size_t(6) << 20;
    const uint32_t limit = uint32_t(std::max<size_t>(1, std::min<size_t>(8, (bytes + unit - 1) / unit)));
    const uint32_t want = std::max(1u, std::min(ntiles, limit));
    Chunking c;
    c.ntiles = ntiles;
    const uint32_t S = row_width / 256;
    const bool strip_shape = rows >= 2 && rows <= 64 && row_width % 256 == 0 && S >= 2 &&
                             size_t(rows) * row_width == n && n_pad == n;
    if (strip_shape && want >= S) {
        c.strips = true;
        uint32_t count = 0;
        auto piece = [&](uint32_t s, uint32_t j0, uint32_t j1) {  // rows [j0, j1) of strip s
            c.set[count] = TileSet{j0 * S + s, j1 - j0, 1, S};
            c.fwd_begin[count] = s, c.fwd_end[count] = j1 == rows ? s + 1 : s;
            c.inv_begin[count] = s, c.inv_end[count] = j0 == 0 ? s + 1 : s;
            ++count;
        };
        if (rows >= 4) {
            piece(0, 0, rows / 4);
            piece(0, rows / 4, rows / 2);
            piece(0, rows / 2, rows);
            for (uint32_t s = 1; s + 1 < S; ++s) {
                if (S >= 4 && (s == 1 || s == S - 2)) {
                    piece(s, 0, rows / 2);
                    piece(s, rows / 2, rows);
                } else {
                    piece(s, 0, rows);
                }
            }
            piece(S - 1, 0, rows / 2);
            piece(S - 1, rows / 2, rows / 2 + rows / 4);
            piece(S - 1, rows / 2 + rows / 4, rows);
        } else {
            for (uint32_t s = 0; s < S; ++s) piece(s, 0, rows);
        }
        c.count = count;
        c.local[0] = 0;
        for (uint32_t i = 0; i < count; ++i) c.local[i + 1] = c.local[i] + c.set[i].count * 256;
        return c;
    }
    const uint32_t per = (ntiles + want - 1) / want;
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
    } else if (want == 3) {
        // the last range about half the others: ceil(2 ntiles / 5) tiles each
        const uint32_t big = (2 * ntiles + 4) / 5;
        for (uint32_t t = 0; t < ntiles; t += big) sizes[count++] = std::min(big, ntiles - t);
    } else {
        for (uint32_t t = 0; t < ntiles; t += per) sizes[count++] = std::min(per, ntiles - t);
    }
    c.count = count;
    uint32_t t0 = 0;
    for (uint32_t i = 0; i < count; ++i) {
        c.set[i] = tile_range(t0, sizes[i]);
        c.local[i] = uint32_t(std::min<size_t>(size_t(t0) * 256, n));
        t0 += sizes[i];
    }
    c.local[count] = n;
    return c;
}

}  // namespace pm
