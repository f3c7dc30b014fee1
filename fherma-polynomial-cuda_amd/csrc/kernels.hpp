// Launch wrappers for the gfx950 kernels of the negacyclic multiplier.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "chunks.hpp"
#include "plan.hpp"
#include "polymul.h"

namespace pm {

// Device views of the plan's tables.
struct DeviceTables {
    const PrimeConst* consts = nullptr;   // [r]
    const int8_t* wconv = nullptr;        // [2] x tile_256x64([conv_rows][conv_k])
    const uint2* tw_fwd = nullptr;        // [r][n2]
    const uint2* tw_inv = nullptr;        // [r][n2]
    const int8_t* wcrt = nullptr;         // tile_256x64([crt_rows][crt_k])
    const uint32_t* q = nullptr;          // [L]
    const uint32_t* neg_p = nullptr;      // [L]
    const uint32_t* corr_t = nullptr;     // [t_words]
    const uint32_t* corr_e = nullptr;     // [4]
};

// Working buffers (row stride n_pad).
struct DeviceWork {
    int8_t* packed = nullptr;    // [2][n_pad/256][conv_k/64][256][64] signed operand bytes
    uint32_t* planes = nullptr;  // [2][r_pad][n_pad] residues of a and b
    int8_t* rtiles = nullptr;    // CRT operand tiles of v_k (aliases the b plane, dead after the row stages)
    int64_t* words = nullptr;    // CRT word sums: [crt_words][n_pad], or [n_pad][ws] (words_by_coefficient)
    uint32_t* t_words = nullptr; // [t_words][n_pad] T scratch of the lane-per-coefficient reconstruction
                                 // (aliases the a plane, dead after the CRT GEMM)
};

// Layout of the CRT word sums.  Up to 62 limbs (t_words <= 64) one lane
// reconstructs one coefficient and the sums are word-major [crt_words][n_pad];
// wider q takes one wave per coefficient over coefficient-major sums
// [n_pad][ws], ws = crt_words rounded up to 16 (128-byte rows).
inline bool words_by_coefficient(const Shape& s) { return s.t_words > 64; }
inline uint32_t word_stride(const Shape& s) { return words_by_coefficient(s) ? (s.crt_words + 15) / 16 * 16 : s.n_pad; }
inline size_t word_sum_elements(const Shape& s) {
    return words_by_coefficient(s) ? size_t(s.n_pad) * word_stride(s) : size_t(s.crt_words) * s.n_pad;
}
// Rows of the a plane the T scratch needs (the lane-per-coefficient reconstruction only).
inline uint32_t scratch_rows(const Shape& s) { return words_by_coefficient(s) ? 0 : s.t_words; }

// Static device workspace of pm_multiply_device_scratch: kScratchBytes per GPU,
// held from the library's first plan on that GPU until the process ends.
constexpr size_t kScratchBytes = PM_SCRATCH_BYTES;
// Looks up the calling thread's current device's copy of the array (once per
// device); a failure is recorded for scratch_address to report.
void init_scratch();
// The device address of `device`'s copy, or nullptr with the reason in *why.
void* scratch_address(int device, std::string* why);
// Working buffers of one device product carved from `base` in 256-byte units
// (base == nullptr: size only); returns the bytes they span.
size_t carve_work(const Shape& s, uint8_t* base, DeviceWork* w);
size_t work_bytes(const Shape& s);
size_t packed_operand_bytes(const Shape& s);

// Stage launches, all on `stream`.  A TileSet names the 256-coefficient tiles a
// launch covers (every tile of it must be one of the plan's, or the launch
// throws std::invalid_argument); all_tiles(s) is every tile.
inline TileSet all_tiles(const Shape& s) { return tile_range(0, s.n_pad / 256); }
// Tiles of `ts` of operands op0 .. op0+nops-1 (0 = a, 1 = b; operand b uses its
// own weights and offsets).  pack reads the k-th tile's rows from rows
// k*256 .. k*256+255 of a and b (for all_tiles: [N][L] as given).
void launch_pack(const Shape& s, const uint32_t* a, const uint32_t* b, int8_t* packed, const TileSet& ts,
                 uint32_t op0, uint32_t nops, hipStream_t stream);
void launch_conversion(const Shape& s, const DeviceTables& t, const int8_t* packed, uint32_t* planes,
                       const TileSet& ts, uint32_t op0, uint32_t nops, hipStream_t stream);
// Forward row-spanning (column) stages of column strips [strip_begin,
// strip_end) of 256 columns each, or of every column with the defaults; a strip
// range needs the register-resident kernel (row_width a multiple of 256, 2 to
// 64 rows).
void launch_ntt_forward_columns(const Shape& s, const DeviceTables& t, uint32_t* planes, uint32_t op0,
                                uint32_t nops, hipStream_t stream, uint32_t strip_begin = 0,
                                uint32_t strip_end = ~0u);
// Reads a's planes from `planes` and b's from `bplanes` ([r_pad][n_pad] each);
// the product lands in a's planes.
void launch_ntt_rows_product(const Shape& s, const DeviceTables& t, uint32_t* planes, const uint32_t* bplanes,
                             hipStream_t stream);
// Inverse row-spanning stages of plane 0 for column strips [strip_begin,
// strip_end) (defaults: every column); writes v_k into the CRT operand tiles
// [n_pad/256][r_pad/16][256][64] (signed bytes) of those columns.
void launch_ntt_inverse_columns(const Shape& s, const DeviceTables& t, uint32_t* planes, int8_t* rtiles,
                                hipStream_t stream, uint32_t strip_begin = 0, uint32_t strip_end = ~0u);
void launch_crt(const Shape& s, const DeviceTables& t, const int8_t* rtiles, int64_t* words, const TileSet& ts,
                hipStream_t stream);
// Reconstructs the coefficients of the tiles of `ts`: output row k*256 + u of c
// (L words each) is coefficient tile_at(ts, k)*256 + u; rows of coefficients at
// or past N are not written.  When reconstruct_writes_once(s), each word of c
// is written once, so c may well be pinned host memory; otherwise (q wider than
// 62 limbs) a coefficient whose quotient estimate was off is written again by
// a later pass.
inline bool reconstruct_writes_once(const Shape& s) { return s.t_words <= 64; }
void launch_reconstruct(const Shape& s, const DeviceTables& t, const int64_t* words, uint32_t* t_scratch,
                        uint32_t* c, const TileSet& ts, hipStream_t stream);
void launch_zero(uint32_t* c, size_t words, hipStream_t stream);

// Grouped launches: every stage of a batch of device products (each its own
// plan: shape, q, tables and working buffers) as one launch per stage and
// kernel variant, instead of one launch per stage and product.  A launch walks
// a work list: entry j owns blocks [start[j], start[j+1]) of the grid and runs
// them for product item[j] with that product's parameters.
struct GroupItem {
    uint32_t n, n_pad, limbs, r, r_pad, conv_rows, conv_k, crt_rows, crt_k, crt_words, ws, t_words;
    uint32_t n1, n2, rows, row_width, s_shift, k_shift, top_m, top_e;
    uint64_t mu;
    const PrimeConst* consts;
    const int8_t* wconv;
    const uint2* tw_fwd;
    const uint2* tw_inv;
    const int8_t* wcrt;
    const uint32_t* q;
    const uint32_t* neg_p;
    const uint32_t* corr_t;
    const uint32_t* corr_e;
    const uint32_t* a;
    const uint32_t* b;
    uint32_t* c;
    int8_t* packed;
    uint32_t* planes;
    int8_t* rtiles;
    int64_t* words;
    uint32_t* t_scratch;
};
// The item of one plain product (a, b -> c on the working buffers w).
GroupItem group_item(const Shape& s, const DeviceTables& t, const DeviceWork& w, const uint32_t* a,
                     const uint32_t* b, uint32_t* c);
// Whether every stage of the shape has a grouped kernel (else the product runs
// on its own): q > 1, and column stages held in registers.
bool groupable(const Shape& s);
// Bytes of device memory a batch of `count` products needs: the items, then
// the launches' work lists.
size_t group_bytes(size_t count);
// Runs the stages of every item, in order, on `stream`, after copying the items
// and the work lists (one copy) into `dev_buffer` (group_bytes(items.size())
// bytes, unused by any launch still queued).
void launch_group(const std::vector<GroupItem>& items, void* dev_buffer, hipStream_t stream);

// Plan-time table construction on the device (plans built with Tables::kDevice):
// beta_k, the merged-twist twiddles, both conversion weight matrices and their
// signed-byte offsets (written into consts[k].off_a / off_b), and the CRT weight
// matrix, each in the layout the product kernels read.  Bit-identical to the
// host construction in plan.cpp.
struct TableBuild {
    PrimeConst* consts = nullptr;        // [r]: p, pinv set; offsets filled in
    const uint32_t* psi = nullptr;       // [r][2]: psi, psi^-1
    const uint64_t* rcp = nullptr;       // [r]: floor((2^64-1)/p_k)
    const uint32_t* m_mod_q = nullptr;   // [r][L]: M_k = P_k mod q
    uint32_t* beta = nullptr;            // [r] out
    int8_t* wconv = nullptr;             // [2] x tile_256x64([conv_rows][conv_k]) out
    uint32_t* tw_fwd = nullptr;          // [r][n2][2] out
    uint32_t* tw_inv = nullptr;          // [r][n2][2] out
    int8_t* wcrt = nullptr;              // tile_256x64([crt_rows][crt_k]) out
};
void launch_build_tables(const Shape& s, const TableBuild& t, hipStream_t stream);

}  // namespace pm
