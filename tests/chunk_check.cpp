// The host pipeline's coefficient chunks (csrc/chunks.hpp) for every
// power-of-two N up to 2^24 and every limb count up to 8192 (beyond the
// 16383-prime cap), with the transform's [rows][row_width] split of the plan
// (row_width = min(N, 2048)): every tile of every chunk lies inside
// [0, ntiles), the chunks cover every tile exactly once, no chunk is empty, at
// most kMaxChunks are used, the runs of each chunk cover [0, N) exactly once
// and fill the chunk-major order without gaps, and with column strips each
// strip's forward column stage is issued once, at the chunk that completes
// the strip, and its inverse stage once, at the first chunk that outputs any
// of its tiles.  Prints "ok <cases> <strip cases>".
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "chunks.hpp"

int main() {
    unsigned long cases = 0, strip_cases = 0;
    for (unsigned log_n = 0; log_n <= 24; ++log_n) {
        const uint32_t n = 1u << log_n, n_pad = (n + 255) / 256 * 256;
        const uint32_t row_width = n < 2048 ? n : 2048, rows = n / row_width;
        const uint32_t S = row_width / 256;
        for (uint32_t limbs = 1; limbs <= 8192; ++limbs) {
            const pm::Chunking c = pm::make_chunking(n_pad, n, limbs, row_width, rows);
            auto fail = [&](const char* what, uint32_t ch) {
                std::printf("%s: N=%u L=%u chunk %u/%u strips=%d\n", what, n, limbs, ch, c.count, int(c.strips));
                std::exit(1);
            };
            if (c.count < 1 || c.count > uint32_t(pm::kMaxChunks) || c.ntiles != n_pad / 256) fail("bad count", 0);
            if (c.local[0] != 0 || c.local[c.count] != n) fail("bad chunk-major extent", 0);
            // tiles and rows are marked one by one only for strip chunkings (at most
            // 64 x 8 tiles); consecutive tile ranges are checked as intervals
            std::vector<uint8_t> tile_seen(c.strips ? c.ntiles : 0, 0);
            std::vector<uint8_t> row_seen(c.strips ? n : 0, 0);
            uint32_t next_tile = 0;
            size_t next_row = 0;
            std::vector<uint32_t> strip_tiles_in(S ? S : 1, 0);
            std::vector<uint8_t> fwd_done(S ? S : 1, 0), inv_done(S ? S : 1, 0);
            for (uint32_t ch = 0; ch < c.count; ++ch) {
                const pm::TileSet& s = c.set[ch];
                if (s.count == 0 || s.w == 0 || s.count % s.w != 0) fail("empty or ragged chunk", ch);
                if (c.runs(ch) > pm::kMaxRuns) fail("more runs than kMaxRuns", ch);
                if (pm::last_tile(s) >= c.ntiles) fail("tile past the end", ch);
                if (c.strips) {
                    for (uint32_t k = 0; k < s.count; ++k) {
                        const uint32_t t = pm::tile_at(s, k);
                        if (t >= c.ntiles || tile_seen[t]) fail("tile outside or repeated", ch);
                        tile_seen[t] = 1;
                    }
                } else {
                    if (s.t0 != next_tile) fail("tile ranges out of order", ch);
                    next_tile = s.t0 + s.count;
                }
                if (c.local[ch + 1] <= c.local[ch]) fail("chunk without rows", ch);
                size_t filled = 0;
                for (uint32_t g = 0; g < c.runs(ch); ++g) {
                    size_t global0, local0, len;
                    c.run(ch, g, n, global0, local0, len);
                    if (local0 != c.local[ch] + filled) fail("runs leave a gap in the chunk-major order", ch);
                    if (global0 + len > n) fail("run past N", ch);
                    if (c.strips) {
                        for (size_t i = global0; i < global0 + len; ++i) {
                            if (row_seen[i]) fail("row repeated", ch);
                            row_seen[i] = 1;
                        }
                    } else {
                        if (global0 != next_row || len == 0) fail("row ranges out of order", ch);
                        next_row = global0 + len;
                    }
                    filled += len;
                }
                if (filled != c.rows_of(ch)) fail("runs do not fill the chunk", ch);
                if (!c.strips) {
                    if (s.w != s.count) fail("a tile-range chunk with several runs", ch);
                    continue;
                }
                // inverse stages first (before this chunk's output), then the tiles, then forward stages
                for (uint32_t st = c.inv_begin[ch]; st < c.inv_end[ch]; ++st) {
                    if (st >= S || inv_done[st] || strip_tiles_in[st] != 0) fail("inverse stage misplaced", ch);
                    inv_done[st] = 1;
                }
                for (uint32_t k = 0; k < s.count; ++k) {
                    const uint32_t st = pm::tile_at(s, k) % S;
                    if (!inv_done[st]) fail("output before its strip's inverse stage", ch);
                    ++strip_tiles_in[st];
                }
                for (uint32_t st = c.fwd_begin[ch]; st < c.fwd_end[ch]; ++st) {
                    if (st >= S || fwd_done[st] || strip_tiles_in[st] != rows) fail("forward stage misplaced", ch);
                    fwd_done[st] = 1;
                }
            }
            if (!c.strips && (next_tile != c.ntiles || next_row != n)) fail("chunks do not cover", 0);
            for (uint8_t seen : tile_seen)
                if (!seen) fail("tile not covered", 0);
            for (uint8_t seen : row_seen)
                if (!seen) fail("row not covered", 0);
            if (c.strips) {
                for (uint32_t st = 0; st < S; ++st)
                    if (!fwd_done[st] || !inv_done[st]) fail("strip stage missing", 0);
                ++strip_cases;
            }
            ++cases;
        }
    }
    std::printf("ok %lu %lu\n", cases, strip_cases);
    return 0;
}
