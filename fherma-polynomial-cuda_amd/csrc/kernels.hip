// gfx950 kernels of the exact negacyclic multiplier c = a*b mod (X^N+1, q).
//
// Pipeline (per product; every stage is integer-only):
//   0. pack_operands        limbs -> 64-byte-aligned K tiles of signed bytes
//   1. conversion_gemm      limb bytes x signed byte weights on INT8 MFMA
//                           (v_mfma_i32_16x16x64_i8), REDC -> residues mod p_k
//                           [reference: tensor_product<true>, solve.cu:143-250]
//   2. ntt_columns_forward  merged-twist negacyclic NTT stages whose butterfly
//                           distance spans whole rows  [high_stages, solve.cu:280-382]
//   3. ntt_rows_product     remaining forward stages, base-case product mod
//                           (X^n1 - zeta), inverse stages, all inside one row
//                           [middle_stages*, solve.cu:385-570]
//   4. ntt_columns_inverse  inverse row-spanning stages -> v_k = x P_k^-1 mod p_k
//                           [high_stages<true>, solve.cu:317-382]
//   5. crt_gemm             residue bytes x signed digits of M_k and of
//                           floor((2^64-1)/p_k) on INT8 MFMA -> word sums of
//                           T = sum v_k M_k and of the quotient sum E
//                           [tensor_product<false>, solve.cu:143-250; CFS:1872-1881]
//   6. reconstruct          K = round(E/2^64), T += K*((-P) mod q), carries,
//                           Barrett reduction by an arbitrary q
//                           [reconstruct_direct, solve.cu:667-716, generalized]
//
// Signed INT8 MFMA: data bytes u in [0,255] enter as u-128 (XOR 0x80); the
// weights are signed base-256 digits; the constant 128*sum(weights) is folded
// into per-prime offsets (conversion) and a fixed big-integer correction (CRT).
//
// Planes, word sums and scratch use a row stride of n_pad = N rounded up to
// 256, so every GEMM tile load is in bounds and 16-byte aligned.
#include "kernels.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

namespace pm {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef uint32_t v2u __attribute__((ext_vector_type(2)));

// GEMM block tile: 256 weight rows x 256 coefficients x 64 bytes of K, eight
// waves in a 4x2 grid, each owning a 64x128 slab = 4x8 MFMA 16x16 tiles
// (128 int32 accumulators per lane; two waves per SIMD).
constexpr uint32_t kTile = 256;
constexpr uint32_t kTK = 64;
constexpr uint32_t kTileBytes = kTile * kTK;  // one operand K step, 16 KB
constexpr uint32_t kGemmThreads = 512;
constexpr uint32_t kLoads = kTileBytes / 16 / kGemmThreads;  // 16-byte loads per thread per operand
constexpr uint32_t kWaveM = 4, kWaveN = 8;    // 16x16 tiles per wave
constexpr uint32_t kWavesN = 2;               // waves across the coefficient dimension
constexpr uint32_t kWaveRows = kWaveM * 16, kWaveCols = kWaveN * 16;
constexpr uint32_t kGroupM = 8;               // raster group height (tiles)

void check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

__device__ __forceinline__ uint32_t add_mod(uint32_t a, uint32_t b, uint32_t p) {
    uint32_t t = a + b;
    return t >= p ? t - p : t;
}

__device__ __forceinline__ uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t p) {
    return a >= b ? a - b : a + p - b;
}

// Shoup: wq = floor(w*2^32/p); for a < 2^32, a*w - hi(a*wq)*p lies in [0, 2p).
__device__ __forceinline__ uint32_t mul_shoup(uint32_t a, uint32_t w, uint32_t wq, uint32_t p) {
    uint32_t q = __umulhi(a, wq);
    uint32_t r = a * w - q * p;
    return r >= p ? r - p : r;
}

// Subtractive Montgomery REDC of t < p*2^32: (t - m p)/2^32, m = t p^-1 mod 2^32.
__device__ __forceinline__ uint32_t redc(uint64_t t, uint32_t p, uint32_t pinv) {
    uint32_t m = uint32_t(t) * pinv;
    uint32_t mh = __umulhi(m, p);
    uint32_t hi = uint32_t(t >> 32);
    uint32_t r = hi - mh;
    return hi < mh ? r + p : r;
}

// REDC for t < 2p*2^32 (four products of residues: 4p^2 < 2p*2^32 as 2p < 2^32):
// hi(t) - hi(m*p) lies in (-p, 2p); one more conditional subtraction than redc.
__device__ __forceinline__ uint32_t redc_wide(uint64_t t, uint32_t p, uint32_t pinv) {
    const uint32_t r = redc(t, p, pinv);  // in [0, 2p) for such t
    return r >= p ? r - p : r;
}

__device__ __forceinline__ uint32_t mont(uint32_t a, uint32_t b, uint32_t p, uint32_t pinv) {
    return redc(uint64_t(a) * b, p, pinv);
}

// 256, 65536 and 2^24 in scalar registers the compiler cannot see through.
// With literal scales each int64 term becomes a sign extension and a 64-bit
// shift and add (about four instructions); with a register scale it is one
// v_mad_i64_i32 (32x32 signed product plus 64-bit addend).
struct DigitScales {
    int32_t s8, s16, s24;
};

__device__ __forceinline__ DigitScales digit_scales() {
    DigitScales s;
    asm volatile("s_mov_b32 %0, 0x100" : "=s"(s.s8));
    asm volatile("s_mov_b32 %0, 0x10000" : "=s"(s.s16));
    asm volatile("s_mov_b32 %0, 0x1000000" : "=s"(s.s24));
    return s;
}

// base + sum_d 256^d v[d], exact in int64.
__device__ __forceinline__ int64_t combine4(const v4i& v, const DigitScales& s, int64_t base = 0) {
    int64_t t = int64_t(v[0]) + base;
    t += int64_t(v[1]) * int64_t(s.s8);
    t += int64_t(v[2]) * int64_t(s.s16);
    t += int64_t(v[3]) * int64_t(s.s24);
    return t;
}

// LDS operand tile: 256 rows of 64 bytes; the four 16-byte chunks of a row are
// XOR-swizzled by row/2.  A 128-bit LDS read serves 8 lanes (128 bytes) per
// pass; a fragment read gives those 8 lanes rows r..r+7 of one chunk, and with
// the swizzle their 16-byte slots (row*4 + chunk') mod 8 are all distinct.  The
// direct-to-LDS loads fill the slots linearly (64 consecutive slots per wave).
__device__ __forceinline__ uint32_t swz(uint32_t row, uint32_t chunk) {
    return row * kTK + (((chunk ^ (row >> 1)) & 3u) << 4);
}

// Block -> (m tile, n tile).  Workgroups are dispatched to the 8 XCDs round
// robin, so first regroup ids so that each XCD owns a contiguous range, then
// walk that range in column-major groups of kGroupM tile rows: the blocks an
// XCD runs together share A tiles (across a column) and B tiles (across a
// group row) in that XCD's L2.
__device__ __forceinline__ void tile_coords(uint32_t bid, uint32_t num_m, uint32_t num_n, uint32_t& tm,
                                            uint32_t& tn) {
    const uint32_t total = num_m * num_n;
    const uint32_t per = total / 8;
    const uint32_t logical = bid < per * 8 ? (bid % 8) * per + bid / 8 : bid;
    const uint32_t in_group = kGroupM * num_n;
    const uint32_t group = logical / in_group, first_m = group * kGroupM;
    const uint32_t height = num_m - first_m < kGroupM ? num_m - first_m : kGroupM;
    const uint32_t local = logical % in_group;
    tm = first_m + local % height;
    tn = local / height;
}

// The K loop, for a block tile of 256 weight rows (A, the pre-tiled weight
// panel [k steps][256][64]) x 256 coefficients (B, the pre-tiled data panel of
// the same layout), 64 bytes of K per step, four stages in flight:
//  * A is not staged: each wave loads its own MFMA fragments straight from
//    global memory.  Fragment mi of step ks for lane l is the 16 bytes of row
//    wm*64 + mi*16 + (l&15), chunk l>>4 -- 1 KB contiguous per fragment, one
//    coalesced load.  (Waves wn = 0 and 1 load the same fragments; the second
//    load mostly hits the CU's L1.)
//  * B goes global -> LDS with direct-to-LDS loads (global_load_lds_dwordx4):
//    lane l of a load fills LDS slot base + l, so each lane fetches the global
//    16-byte chunk that the XOR-swizzled tile layout (swz) puts in that slot.
//    No register staging and no ds_write instructions.
//  * Loads run three steps ahead: four LDS stages (64 KB) and four register
//    sets of A fragments rotate.  Step k waits until its own loads have landed
//    (the 12 loads of steps k+1, k+2 may stay in flight: 4 A + 2 B per lane and
//    step) and until this wave's LDS reads are done, then one barrier; after
//    it, no wave still reads stage (k-1)%4, which step k+3's loads refill.
// The barrier is a plain s_barrier with a compiler memory clobber, not
// __syncthreads(): the latter's release fence would wait for every
// outstanding direct-to-LDS load (vmcnt 0) and undo the prefetch.
// Measured on the conversion and CRT shapes of N=65536, L=1747 against the
// two-stage loop that staged both operands through registers and ds_write:
// 2-3 % faster on the conversion GEMM, 5-10 % on the CRT GEMM.
struct GemmStages {
    int8_t* s[4];
};

__device__ __forceinline__ void gemm_mainloop(const int8_t* __restrict__ a_panel, const int8_t* __restrict__ b_panel,
                                              uint32_t ksteps, const GemmStages& st, v4i (&acc)[kWaveM][kWaveN]) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t wm = wave / kWavesN, wn = wave % kWavesN;
    const v4i* ap = reinterpret_cast<const v4i*>(a_panel + (wm * kWaveRows + (lane & 15)) * kTK + (lane >> 4) * 16);
    // B: slot s = tid + 512 j of a stage holds row s>>2, swizzled chunk s&3,
    // i.e. the row's chunk (s&3) ^ ((row>>1)&3) in the global tile.
    uint32_t boff[kLoads];
#pragma unroll
    for (uint32_t j = 0; j < kLoads; ++j) {
        const uint32_t slot = tid + kGemmThreads * j, row = slot >> 2, chunk = (slot & 3) ^ ((row >> 1) & 3);
        boff[j] = row * kTK + chunk * 16;
    }
    auto clamp = [&](uint32_t k) { return k < ksteps ? k : ksteps - 1; };  // past the end: repeat the last step
    auto load_b = [&](uint32_t ks, int8_t* stage) {
        const int8_t* src = b_panel + size_t(ks) * kTileBytes;
#pragma unroll
        for (uint32_t j = 0; j < kLoads; ++j)
            __builtin_amdgcn_global_load_lds(reinterpret_cast<const void*>(src + boff[j]),
                                             (__attribute__((address_space(3))) void*)(stage + (wave * 64 + kGemmThreads * j) * 16),
                                             16, 0, 0);
    };
    auto load_a = [&](uint32_t ks, v4i (&fa)[kWaveM]) {
#pragma unroll
        for (uint32_t mi = 0; mi < kWaveM; ++mi) fa[mi] = ap[size_t(ks) * (kTileBytes / 16) + mi * 64];
    };
    v4i f0[kWaveM], f1[kWaveM], f2[kWaveM], f3[kWaveM], fb[kWaveN];
    auto step = [&](uint32_t k, const int8_t* stage, const v4i (&fa)[kWaveM], int8_t* next_stage,
                    v4i (&next_fa)[kWaveM]) {
        // vmcnt <= 12 (gfx9 encoding: low bits 3:0, high bits 15:14), lgkmcnt 0, expcnt untouched
        constexpr int kWait = (12 & 15) | ((12 >> 4) << 14) | (0x7 << 4);
        __builtin_amdgcn_s_waitcnt(kWait);
        asm volatile("s_barrier" ::: "memory");
        load_b(clamp(k + 3), next_stage);
        load_a(clamp(k + 3), next_fa);
#pragma unroll
        for (uint32_t ni = 0; ni < kWaveN; ++ni)
            fb[ni] = *reinterpret_cast<const v4i*>(stage + swz(wn * kWaveCols + ni * 16 + (lane & 15), lane >> 4));
#pragma unroll
        for (uint32_t mi = 0; mi < kWaveM; ++mi)
#pragma unroll
            for (uint32_t ni = 0; ni < kWaveN; ++ni)
                acc[mi][ni] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[mi], fb[ni], acc[mi][ni], 0, 0, 0);
    };
    load_b(0, st.s[0]);
    load_a(0, f0);
    load_b(clamp(1), st.s[1]);
    load_a(clamp(1), f1);
    load_b(clamp(2), st.s[2]);
    load_a(clamp(2), f2);
    uint32_t k0 = 0;
    for (; k0 + 4 <= ksteps; k0 += 4) {
        step(k0, st.s[0], f0, st.s[3], f3);
        step(k0 + 1, st.s[1], f1, st.s[0], f0);
        step(k0 + 2, st.s[2], f2, st.s[1], f1);
        step(k0 + 3, st.s[3], f3, st.s[2], f2);
    }
    if (k0 < ksteps) step(k0, st.s[0], f0, st.s[3], f3);
    if (k0 + 1 < ksteps) step(k0 + 1, st.s[1], f1, st.s[0], f0);
    if (k0 + 2 < ksteps) step(k0 + 2, st.s[2], f2, st.s[1], f1);
    // No direct-to-LDS load may outlive the block's LDS (the loads past the
    // last step refill stages nobody reads).
    __builtin_amdgcn_s_waitcnt((0x7 << 4) | (0xF << 8));  // vmcnt 0
}

// The four 16 KB LDS stages of gemm_mainloop, as separate arrays of the kernel
// (so the compiler can tell a stage being filled from the one being read).
#define PM_GEMM_STAGES(name)                                                                               \
    __shared__ __attribute__((aligned(16))) int8_t name##0[kTileBytes], name##1[kTileBytes],                \
        name##2[kTileBytes], name##3[kTileBytes];                                                          \
    const GemmStages name{{name##0, name##1, name##2, name##3}}

// ---------------------------------------------------------------------------
// 0. operand packing: [N][L] u32 -> [2][n_pad/256][conv_k/64][256][64] bytes,
//    each byte XOR 0x80 (u -> u-128 as int8); padding words (beyond L, or rows
//    beyond N) hold the value 0.  One wave per coefficient row (two or four
//    short rows per wave, below): lane l moves words l, l+64, ... of the row,
//    so every load reads 256 contiguous bytes and each 128-byte line of the
//    input is consumed by one wave (a block holds consecutive rows, so only
//    lines at block edges are shared).
// ---------------------------------------------------------------------------
// Rows of one wave: four rows of 16 lanes each, so that a store instruction
// writes four consecutive coefficients' 64-byte K-step chunks, 256 contiguous
// bytes of the packed tile (with one row per wave, it wrote four chunks 16 KB
// apart: 0.391 against 0.370 ms at N = 65536 with 1747 limbs).  Rows of two K
// steps (L of 9 to 16 limbs) go two at a time: 32 lanes cover their 32 words.
constexpr uint32_t pack_rows_per_wave(uint32_t ksteps) { return ksteps == 2 ? 2 : 4; }

// Packs the tiles of `ts` of operands op0 .. op0+nops-1.  The k-th tile's rows
// are read from rows k*256 .. k*256+255 of a (or b): the caller passes the
// operands' rows in the set's order (all of [N][L] for the set of every tile).
__device__ __forceinline__ void pack_body(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                                          uint32_t* __restrict__ packed, uint32_t n, uint32_t limbs,
                                          uint32_t ksteps, uint32_t ntiles, TileSet ts, uint32_t op0,
                                          uint32_t nops, uint32_t bx) {
    const uint32_t rpw = pack_rows_per_wave(ksteps), span = 64 / rpw;  // lanes per row
    const uint32_t lane = (threadIdx.x & 63) % span;
    const uint32_t per_op = ts.count * kTile;
    // row over all operands of the launch
    const uint32_t gr = (bx * 4 + (threadIdx.x >> 6)) * rpw + (threadIdx.x & 63) / span;
    if (gr >= nops * per_op) return;
    const uint32_t op = op0 + gr / per_op, rr = gr % per_op;
    const uint32_t tn = tile_at(ts, rr / kTile), row = rr % kTile, coef = tn * kTile + row;
    const bool live = coef < n;
    // Rows past n read (and discard) the set's first row, so every load is
    // unconditional: kPackLoads loads per lane are in flight before their
    // stores, instead of one load, a wait and a store at a time.
    const uint32_t* src = (op ? b : a) + size_t(live ? rr : 0) * limbs;
    uint32_t* dst = packed + (size_t(op) * ntiles + tn) * ksteps * (kTileBytes / 4) + row * (kTK / 4);
    const uint32_t words = ksteps * (kTK / 4);
    constexpr uint32_t kPackLoads = 8;
    for (uint32_t w0 = lane; w0 < words; w0 += span * kPackLoads) {
        uint32_t v[kPackLoads];
#pragma unroll
        for (uint32_t u = 0; u < kPackLoads; ++u) {
            const uint32_t w = w0 + span * u;
            v[u] = src[w < limbs ? w : limbs - 1];  // clamped into the row; padding words are zeroed below
        }
#pragma unroll
        for (uint32_t u = 0; u < kPackLoads; ++u) {
            const uint32_t w = w0 + span * u;
            if (w < words)
                dst[size_t(w / (kTK / 4)) * (kTileBytes / 4) + w % (kTK / 4)] =
                    ((live && w < limbs) ? v[u] : 0u) ^ 0x80808080u;
        }
    }
}

__global__ __launch_bounds__(256) void pack_operands(const uint32_t* __restrict__ a,
                                                     const uint32_t* __restrict__ b,
                                                     uint32_t* __restrict__ packed, uint32_t n, uint32_t limbs,
                                                     uint32_t ksteps, uint32_t ntiles, TileSet ts, uint32_t op0,
                                                     uint32_t nops) {
    pack_body(a, b, packed, n, limbs, ksteps, ntiles, ts, op0, nops, blockIdx.x);
}

// ---------------------------------------------------------------------------
// 1. conversion GEMM: G[4k+d][i] = sum_j digit_d(lambda*256^j mod p_k) * (byte_j(x_i)-128)
//    then x_i*lambda mod p_k = REDC(sum_d 256^d G[4k+d][i] + off_k).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void conversion_body(const int8_t* __restrict__ wtiles, const int8_t* __restrict__ packed,
                                                const PrimeConst* __restrict__ consts, uint32_t* __restrict__ planes,
                                                uint32_t n, uint32_t n_pad, uint32_t r, uint32_t r_pad,
                                                uint32_t conv_rows, uint32_t ksteps, TileSet ts, uint32_t op0,
                                                uint32_t nops, uint32_t bid) {
    PM_GEMM_STAGES(stages);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t wm = wave / kWavesN, wn = wave % kWavesN;
    const uint32_t ntiles = n_pad / kTile, num_m = conv_rows / kTile;
    uint32_t tm, tn;
    tile_coords(bid, num_m, nops * ts.count, tm, tn);  // the set's coefficient tiles of nops operands
    const uint32_t op = op0 + tn / ts.count, tcol = tile_at(ts, tn % ts.count);
    const int8_t* a_panel = wtiles + (size_t(op) * num_m + tm) * ksteps * kTileBytes;
    const int8_t* b_panel = packed + (size_t(op) * ntiles + tcol) * ksteps * kTileBytes;

    v4i acc[kWaveM][kWaveN];
#pragma unroll
    for (uint32_t mi = 0; mi < kWaveM; ++mi)
#pragma unroll
        for (uint32_t ni = 0; ni < kWaveN; ++ni) acc[mi][ni] = v4i{0, 0, 0, 0};
    gemm_mainloop(a_panel, b_panel, ksteps, stages, acc);

    // Lane holds rows 4*(lane/16)+j (j = digit) of each 16-row tile: one prime.
    // The digit sum is sum_j c_j*(byte_j - 128) with |c_j| <= p/2 over 4L bytes,
    // so |sum| <= 256*L*p; adding p*2^31 (0 mod p) to off_k makes t non-negative
    // and below p*2^32, the range of the unsigned REDC, while 256*L + 1 <= 2^31.
    // Plans allow at most 16383 primes of < 32 bits covering (q-1)^2, so
    // q < 2^262144 and L <= 8192: 256*L <= 2^21.
    const DigitScales ds = digit_scales();
    const uint32_t m0 = tm * kTile + wm * kWaveRows, c0 = tcol * kTile + wn * kWaveCols;
#pragma unroll
    for (uint32_t mi = 0; mi < kWaveM; ++mi) {
        const uint32_t kp = (m0 + mi * 16) / 4 + (lane >> 4);
        if (kp >= r) continue;
        const PrimeConst pc = consts[kp];
        const int64_t off = int64_t(op ? pc.off_b : pc.off_a) + (int64_t(pc.p) << 31);
        uint32_t* out = planes + (size_t(op) * r_pad + kp) * n_pad;
#pragma unroll
        for (uint32_t ni = 0; ni < kWaveN; ++ni) {
            const uint32_t coef = c0 + ni * 16 + (lane & 15);
            if (coef < n) out[coef] = redc(uint64_t(combine4(acc[mi][ni], ds, off)), pc.p, pc.pinv);
        }
    }
}

__global__ __launch_bounds__(512) void conversion_gemm(const int8_t* __restrict__ wtiles,
                                                       const int8_t* __restrict__ packed,
                                                       const PrimeConst* __restrict__ consts,
                                                       uint32_t* __restrict__ planes, uint32_t n, uint32_t n_pad,
                                                       uint32_t r, uint32_t r_pad, uint32_t conv_rows,
                                                       uint32_t ksteps, TileSet ts, uint32_t op0, uint32_t nops) {
    conversion_body(wtiles, packed, consts, planes, n, n_pad, r, r_pad, conv_rows, ksteps, ts, op0, nops, blockIdx.x);
}

// ---------------------------------------------------------------------------
// 5. CRT GEMM: G[c][i] = sum_{k,d} w[c][4k+d] * (byte_d(v_{k,i}) - 128),
//    D_j(i) = sum_t 256^t G[4j+t][i]   (int64, exact)
//    stored word-major words[j][i] (row stride n_pad), or, for the
//    wave-per-coefficient reconstruction, coefficient-major words[i][j] (row
//    stride ws): a wave's 4 x 16 words of 16 coefficients are then 128-byte
//    runs per coefficient, written one after another.
// ---------------------------------------------------------------------------
template <bool kByCoefficient>
__device__ __forceinline__ void crt_body(const int8_t* __restrict__ wtiles, const int8_t* __restrict__ rtiles,
                                         int64_t* __restrict__ words, uint32_t n, uint32_t n_pad, uint32_t crt_rows,
                                         uint32_t ksteps, uint32_t crt_words, uint32_t ws, TileSet ts,
                                         uint32_t bid) {
    PM_GEMM_STAGES(stages);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t wm = wave / kWavesN, wn = wave % kWavesN;
    const uint32_t num_m = crt_rows / kTile;
    uint32_t tm, tn;
    tile_coords(bid, num_m, ts.count, tm, tn);  // the set's coefficient tiles
    tn = tile_at(ts, tn);
    const int8_t* a_panel = wtiles + size_t(tm) * ksteps * kTileBytes;
    const int8_t* b_panel = rtiles + size_t(tn) * ksteps * kTileBytes;

    v4i acc[kWaveM][kWaveN];
#pragma unroll
    for (uint32_t mi = 0; mi < kWaveM; ++mi)
#pragma unroll
        for (uint32_t ni = 0; ni < kWaveN; ++ni) acc[mi][ni] = v4i{0, 0, 0, 0};
    gemm_mainloop(a_panel, b_panel, ksteps, stages, acc);

    const DigitScales ds = digit_scales();
    const uint32_t m0 = tm * kTile + wm * kWaveRows, c0 = tn * kTile + wn * kWaveCols;
    if constexpr (kByCoefficient) {
#pragma unroll
        for (uint32_t ni = 0; ni < kWaveN; ++ni) {
            const uint32_t coef = c0 + ni * 16 + (lane & 15);
            int64_t* out = words + size_t(coef) * ws;
#pragma unroll
            for (uint32_t mi = 0; mi < kWaveM; ++mi) {
                const uint32_t word = (m0 + mi * 16) / 4 + (lane >> 4);
                if (word < crt_words && coef < n) out[word] = combine4(acc[mi][ni], ds);
            }
        }
    } else {
#pragma unroll
        for (uint32_t mi = 0; mi < kWaveM; ++mi) {
            const uint32_t word = (m0 + mi * 16) / 4 + (lane >> 4);
            if (word >= crt_words) continue;
            int64_t* out = words + size_t(word) * n_pad;
#pragma unroll
            for (uint32_t ni = 0; ni < kWaveN; ++ni) {
                const uint32_t coef = c0 + ni * 16 + (lane & 15);
                if (coef < n) out[coef] = combine4(acc[mi][ni], ds);
            }
        }
    }
}

template <bool kByCoefficient>
__global__ __launch_bounds__(512) void crt_gemm(const int8_t* __restrict__ wtiles,
                                                const int8_t* __restrict__ rtiles,
                                                int64_t* __restrict__ words, uint32_t n, uint32_t n_pad,
                                                uint32_t crt_rows, uint32_t ksteps, uint32_t crt_words,
                                                uint32_t ws, TileSet ts) {
    crt_body<kByCoefficient>(wtiles, rtiles, words, n, n_pad, crt_rows, ksteps, crt_words, ws, ts, blockIdx.x);
}

// ---------------------------------------------------------------------------
// 2-4. NTT.  Flat index i = row*C + col.  Forward: Cooley-Tukey with the
// negacyclic twist merged into psi^bitrev twiddles, m = 1, 2, .., n2/2 groups
// at flat distance N/(2m) down to n1.  Inverse: Gentleman-Sande, mirror order.
// Plane rows have stride ps (= n_pad).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void ntt_columns_forward(uint32_t* __restrict__ planes,
                                                           const PrimeConst* __restrict__ consts,
                                                           const uint2* __restrict__ tw, uint32_t ps,
                                                           uint32_t n2, uint32_t C, uint32_t rows,
                                                           uint32_t cw, uint32_t r, uint32_t op0) {
    extern __shared__ uint32_t tile[];  // [rows][cw]
    const uint32_t col0 = blockIdx.x * cw, k = blockIdx.y, op = op0 + blockIdx.z;
    const uint32_t p = consts[k].p;
    uint32_t* data = planes + (size_t(op) * r + k) * ps;
    const uint2* w = tw + size_t(k) * n2;
    const uint32_t total = rows * cw;
    for (uint32_t t = threadIdx.x; t < total; t += blockDim.x)
        tile[t] = data[size_t(t / cw) * C + col0 + t % cw];
    __syncthreads();
    for (uint32_t m = 1, half = rows / 2; m < rows; m *= 2, half /= 2) {
        for (uint32_t t = threadIdx.x; t < total / 2; t += blockDim.x) {
            uint32_t pr = t / cw, c = t % cw;
            uint32_t g = pr / half, jr = pr % half, row = 2 * half * g + jr;
            uint2 tw_g = w[m + g];
            uint32_t u = tile[row * cw + c];
            uint32_t v = mul_shoup(tile[(row + half) * cw + c], tw_g.x, tw_g.y, p);
            tile[row * cw + c] = add_mod(u, v, p);
            tile[(row + half) * cw + c] = sub_mod(u, v, p);
        }
        __syncthreads();
    }
    for (uint32_t t = threadIdx.x; t < total; t += blockDim.x)
        data[size_t(t / cw) * C + col0 + t % cw] = tile[t];
}

// Row stages for any row width C <= 2048 and block size n1, through LDS;
// lds32 holds 2*C words.
__device__ __forceinline__ void rows_product_body(uint32_t* __restrict__ planes, const uint32_t* __restrict__ bplanes,
                                                  const PrimeConst* __restrict__ consts,
                                                  const uint2* __restrict__ twf, const uint2* __restrict__ twi,
                                                  uint32_t ps, uint32_t n1, uint32_t n2, uint32_t C, uint32_t rows,
                                                  uint32_t* lds32, uint32_t row, uint32_t k) {
    uint32_t* A = lds32;  // A[C], B[C]
    uint32_t* B = lds32 + C;
    const uint32_t p = consts[k].p, pinv = consts[k].pinv;
    uint32_t* pa = planes + size_t(k) * ps + size_t(row) * C;
    const uint32_t* pb = bplanes + size_t(k) * ps + size_t(row) * C;
    const uint2* wf = twf + size_t(k) * n2;
    const uint2* wi = twi + size_t(k) * n2;
    for (uint32_t t = threadIdx.x; t < C; t += blockDim.x) {
        A[t] = pa[t];
        B[t] = pb[t];
    }
    __syncthreads();
    // forward stages with flat distance < C
    for (uint32_t m = rows, t = C / 2; m < n2; m *= 2, t /= 2) {
        for (uint32_t idx = threadIdx.x; idx < C / 2; idx += blockDim.x) {
            uint32_t gl = idx / t, jr = idx % t, j = 2 * t * gl + jr;
            uint2 w = wf[m + row * (C / (2 * t)) + gl];
            uint32_t u = A[j], v = mul_shoup(A[j + t], w.x, w.y, p);
            A[j] = add_mod(u, v, p);
            A[j + t] = sub_mod(u, v, p);
            u = B[j];
            v = mul_shoup(B[j + t], w.x, w.y, p);
            B[j] = add_mod(u, v, p);
            B[j + t] = sub_mod(u, v, p);
        }
        __syncthreads();
    }
    // base case: block gb holds A, B mod (X^n1 - zeta_gb), zeta_gb = psi^(2 bitrev(gb) + 1)
    uint32_t res[8];
    {
        uint32_t cnt = 0;
        for (uint32_t idx = threadIdx.x; idx < C; idx += blockDim.x, ++cnt) {
            const uint32_t s = idx % n1, blk = idx / n1, gb = row * (C / n1) + blk;
            const uint32_t* ab = A + blk * n1;
            const uint32_t* bb = B + blk * n1;
            uint32_t lo = 0, hi = 0;
            for (uint32_t t = 0; t <= s; ++t) lo = add_mod(lo, mont(ab[t], bb[s - t], p, pinv), p);
            for (uint32_t t = s + 1; t < n1; ++t) hi = add_mod(hi, mont(ab[t], bb[s + n1 - t], p, pinv), p);
            uint32_t value = lo;
            if (n1 > 1) {
                if (n2 == 1) {
                    value = sub_mod(lo, hi, p);  // zeta = -1
                } else {
                    uint2 w = wf[n2 / 2 + gb / 2];  // zeta_{2g} = w, zeta_{2g+1} = -w
                    uint32_t z = mul_shoup(hi, w.x, w.y, p);
                    value = (gb & 1) ? sub_mod(lo, z, p) : add_mod(lo, z, p);
                }
            }
            res[cnt] = value;
        }
    }
    __syncthreads();
    {
        uint32_t cnt = 0;
        for (uint32_t idx = threadIdx.x; idx < C; idx += blockDim.x, ++cnt) A[idx] = res[cnt];
    }
    __syncthreads();
    // inverse stages with flat distance < C
    for (uint32_t m = n2 / 2, t = n1; m >= rows && m > 0; m /= 2, t *= 2) {
        for (uint32_t idx = threadIdx.x; idx < C / 2; idx += blockDim.x) {
            uint32_t gl = idx / t, jr = idx % t, j = 2 * t * gl + jr;
            uint2 w = wi[m + row * (C / (2 * t)) + gl];
            uint32_t u = A[j], v = A[j + t];
            A[j] = add_mod(u, v, p);
            A[j + t] = mul_shoup(sub_mod(u, v, p), w.x, w.y, p);
        }
        __syncthreads();
    }
    for (uint32_t t = threadIdx.x; t < C; t += blockDim.x) pa[t] = A[t];
}

__global__ __launch_bounds__(256) void ntt_rows_product(uint32_t* __restrict__ planes,
                                                        const uint32_t* __restrict__ bplanes,
                                                        const PrimeConst* __restrict__ consts,
                                                        const uint2* __restrict__ twf,
                                                        const uint2* __restrict__ twi, uint32_t ps,
                                                        uint32_t n1, uint32_t n2, uint32_t C,
                                                        uint32_t rows, uint32_t r) {
    extern __shared__ uint32_t lds32[];
    rows_product_body(planes, bplanes, consts, twf, twi, ps, n1, n2, C, rows, lds32, blockIdx.x, blockIdx.y);
}

__global__ __launch_bounds__(256) void ntt_columns_inverse(uint32_t* __restrict__ plane,
                                                           const PrimeConst* __restrict__ consts,
                                                           const uint2* __restrict__ tw, uint32_t ps,
                                                           uint32_t n2, uint32_t C, uint32_t rows,
                                                           uint32_t cw) {
    extern __shared__ uint32_t tile[];  // [rows][cw]
    const uint32_t col0 = blockIdx.x * cw, k = blockIdx.y;
    const uint32_t p = consts[k].p;
    uint32_t* data = plane + size_t(k) * ps;
    const uint2* w = tw + size_t(k) * n2;
    const uint32_t total = rows * cw;
    for (uint32_t t = threadIdx.x; t < total; t += blockDim.x)
        tile[t] = data[size_t(t / cw) * C + col0 + t % cw];
    __syncthreads();
    for (uint32_t m = rows / 2, half = 1; m >= 1; m /= 2, half *= 2) {
        for (uint32_t t = threadIdx.x; t < total / 2; t += blockDim.x) {
            uint32_t pr = t / cw, c = t % cw;
            uint32_t g = pr / half, jr = pr % half, row = 2 * half * g + jr;
            uint2 tw_g = w[m + g];
            uint32_t u = tile[row * cw + c], v = tile[(row + half) * cw + c];
            tile[row * cw + c] = add_mod(u, v, p);
            tile[(row + half) * cw + c] = mul_shoup(sub_mod(u, v, p), tw_g.x, tw_g.y, p);
        }
        __syncthreads();
    }
    for (uint32_t t = threadIdx.x; t < total; t += blockDim.x)
        data[size_t(t / cw) * C + col0 + t % cw] = tile[t];
}

// ---------------------------------------------------------------------------
// Register-resident NTT kernels for the usual shapes (rows <= 64, C = 2048).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void ct_butterfly(uint32_t& a, uint32_t& b, uint2 w, uint32_t p) {
    const uint32_t v = mul_shoup(b, w.x, w.y, p);
    b = sub_mod(a, v, p);
    a = add_mod(a, v, p);
}

__device__ __forceinline__ void gs_butterfly(uint32_t& a, uint32_t& b, uint2 w, uint32_t p) {
    const uint32_t u = a, v = b;
    a = add_mod(u, v, p);
    b = mul_shoup(sub_mod(u, v, p), w.x, w.y, p);
}

// Forward row-spanning stages: one thread per (operand, prime, column) holding
// the column's ROWS values; the stage twiddles are uniform over the block.
// Columns col0 + [0, 256 * gridDim.x).
template <int ROWS>
__device__ __forceinline__ void forward_reg_body(uint32_t* __restrict__ planes, const PrimeConst* __restrict__ consts,
                                                 const uint2* __restrict__ tw, uint32_t ps, uint32_t n2, uint32_t C,
                                                 uint32_t r, uint32_t op0, uint32_t col0, uint32_t bx, uint32_t by,
                                                 uint32_t bz) {
    const uint32_t col = col0 + bx * 256 + threadIdx.x, k = by, op = op0 + bz;
    const uint32_t p = consts[k].p;
    uint32_t* data = planes + (size_t(op) * r + k) * ps + col;
    const uint2* w = tw + size_t(k) * n2;
    uint32_t x[ROWS];
#pragma unroll
    for (int j = 0; j < ROWS; ++j) x[j] = data[size_t(j) * C];
#pragma unroll
    for (int m = 1, half = ROWS / 2; m < ROWS; m *= 2, half /= 2)
#pragma unroll
        for (int g = 0; g < m; ++g) {
            const uint2 t = w[m + g];
#pragma unroll
            for (int jr = 0; jr < half; ++jr) ct_butterfly(x[2 * half * g + jr], x[2 * half * g + jr + half], t, p);
        }
#pragma unroll
    for (int j = 0; j < ROWS; ++j) data[size_t(j) * C] = x[j];
}

template <int ROWS>
__global__ __launch_bounds__(256) void ntt_columns_forward_reg(uint32_t* __restrict__ planes,
                                                               const PrimeConst* __restrict__ consts,
                                                               const uint2* __restrict__ tw, uint32_t ps,
                                                               uint32_t n2, uint32_t C, uint32_t r, uint32_t op0,
                                                               uint32_t col0) {
    forward_reg_body<ROWS>(planes, consts, tw, ps, n2, C, r, op0, col0, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Inverse row-spanning stages for 16 primes x 16 columns per block, then the
// residues v_{k,i} are written, XOR 0x80808080, into the CRT GEMM's operand
// tiles [n_pad/256][r_pad/16][256 coefficients][16 primes x 4 bytes] through
// an LDS transpose, so each coefficient's 64-byte K chunk is one store.
// ROWS = 1 is a pure transpose (no stages).  Column groups g0 + [0, gridDim.x)
// of 16 columns each.
template <int ROWS>
__device__ __forceinline__ void inverse_tiles_body(const uint32_t* __restrict__ plane,
                                                   const PrimeConst* __restrict__ consts,
                                                   const uint2* __restrict__ tw, uint32_t ps, uint32_t n, uint32_t n2,
                                                   uint32_t C, uint32_t r, uint32_t ksteps, int8_t* __restrict__ tiles,
                                                   uint32_t g0, uint32_t bx, uint32_t by, uint32_t nx) {
    // lds[j][column][slot]: the four 16-byte groups of a column's 16 primes are
    // XOR-swizzled by column/2, so the transpose reads each 16-byte store as one
    // aligned 128-bit LDS read (8 lanes of a read pass: two columns, distinct
    // banks) and the 4-byte writes of a wave's 32-lane halves are at most 2-way
    // conflicted (4-way unswizzled).
    __shared__ __attribute__((aligned(16))) uint32_t lds[ROWS][16][16];
    auto slot = [](uint32_t c, uint32_t kk) { return ((((kk >> 2) ^ (c >> 1)) & 3u) << 2) | (kk & 3u); };
    const uint32_t kk = threadIdx.x >> 4, cc = threadIdx.x & 15;
    // A block reads 64 bytes of each row: half a 128-byte line, whose other half
    // is the next column group's.  Workgroups go to the 8 XCDs round robin by
    // linear id, so column groups 2m and 2m+1 are given to blocks 8 ids apart
    // (the same XCD when the grid's x extent is a multiple of 16): the second
    // read of each line hits that XCD's L2 instead of memory.
    const uint32_t cx = (nx & 15u) == 0 ? (bx & ~15u) | ((bx & 7u) << 1) | ((bx >> 3) & 1u) : bx;
    const uint32_t c0 = (g0 + cx) * 16, kg = by, k = kg * 16 + kk;
    uint32_t x[ROWS];
    if (k < r) {
        const uint32_t p = consts[k].p;
        const uint32_t* data = plane + size_t(k) * ps + c0 + cc;
        const uint2* w = tw + size_t(k) * n2;
#pragma unroll
        for (int j = 0; j < ROWS; ++j) x[j] = data[size_t(j) * C];
#pragma unroll
        for (int m = ROWS / 2, half = 1; m >= 1; m /= 2, half *= 2)
#pragma unroll
            for (int g = 0; g < m; ++g) {
                const uint2 t = w[m + g];
#pragma unroll
                for (int jr = 0; jr < half; ++jr) gs_butterfly(x[2 * half * g + jr], x[2 * half * g + jr + half], t, p);
            }
    } else {
#pragma unroll
        for (int j = 0; j < ROWS; ++j) x[j] = 0;
    }
#pragma unroll
    for (int j = 0; j < ROWS; ++j) lds[j][cc][slot(cc, kk)] = x[j] ^ 0x80808080u;
    __syncthreads();
    for (uint32_t idx = threadIdx.x; idx < uint32_t(ROWS) * 64; idx += 256) {
        const uint32_t q = idx & 3, c2 = (idx >> 2) & 15, j = idx >> 6;
        const uint32_t i = j * C + c0 + c2;
        if (i < n) {
            const v4u v = *reinterpret_cast<const v4u*>(&lds[j][c2][slot(c2, 4 * q)]);
            *reinterpret_cast<v4u*>(tiles + ((size_t(i >> 8) * ksteps + kg) * 256 + (i & 255)) * 64 + q * 16) = v;
        }
    }
}

template <int ROWS>
__global__ __launch_bounds__(256) void ntt_columns_inverse_tiles(const uint32_t* __restrict__ plane,
                                                                 const PrimeConst* __restrict__ consts,
                                                                 const uint2* __restrict__ tw, uint32_t ps,
                                                                 uint32_t n, uint32_t n2, uint32_t C, uint32_t r,
                                                                 uint32_t ksteps, int8_t* __restrict__ tiles,
                                                                 uint32_t g0) {
    inverse_tiles_body<ROWS>(plane, consts, tw, ps, n, n2, C, r, ksteps, tiles, g0, blockIdx.x, blockIdx.y, gridDim.x);
}

// Row stages for C = 2048: one block per (row, prime), 256 threads x 8 values
// per operand.  The 11 butterfly levels inside a row are done three at a time
// in registers, with an LDS exchange between groups:
//   layout A: jj = t + 256e      distances 1024, 512, 256
//   layout B: jj = 256(t/32) + (t%32) + 32e   distances 128, 64, 32
//   layout C: jj = 32(t/4) + (t%4) + 4e       distances 16, 8, 4
//   layout D: jj = 8t + e        distances 2, 1, and the base case (n1 <= 8)
// Forward stops at distance n1, the base case multiplies each n1-block mod
// (X^n1 - zeta), and the inverse climbs back from n1 to 1024.
constexpr uint32_t kRowC = 2048;

__device__ __forceinline__ uint32_t layout_index(int layout, uint32_t t, int e) {
    switch (layout) {
        case 0: return t + 256u * e;
        case 1: return 256u * (t >> 5) + (t & 31u) + 32u * e;
        case 2: return 32u * (t >> 2) + (t & 3u) + 4u * e;
        default: return 8u * t + e;
    }
}

// forward levels of one layout: distances stride*2^(levels-1) .. stride (those >= n1)
__device__ __forceinline__ void row_forward_group(uint32_t (&x)[8], int layout, int levels, uint32_t t,
                                                  uint32_t stride, uint32_t n1, uint32_t base_index,
                                                  const uint2* w, uint32_t p) {
#pragma unroll
    for (int b = levels - 1; b >= 0; --b) {
        const uint32_t d = stride << b;
        if (d < n1) continue;
        const uint32_t shift = 31 - __clz(2 * d);  // log2(2d)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            if (e & (1 << b)) continue;
            const uint32_t jj = layout_index(layout, t, e);
            const uint2 tw = w[(base_index >> shift) + (jj >> shift)];
            ct_butterfly(x[e], x[e + (1 << b)], tw, p);
        }
    }
}

__device__ __forceinline__ void row_inverse_group(uint32_t (&x)[8], int layout, int levels, uint32_t t,
                                                  uint32_t stride, uint32_t n1, uint32_t base_index,
                                                  const uint2* w, uint32_t p) {
#pragma unroll
    for (int b = 0; b < levels; ++b) {
        const uint32_t d = stride << b;
        if (d < n1) continue;
        const uint32_t shift = 31 - __clz(2 * d);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            if (e & (1 << b)) continue;
            const uint32_t jj = layout_index(layout, t, e);
            const uint2 tw = w[(base_index >> shift) + (jj >> shift)];
            gs_butterfly(x[e], x[e + (1 << b)], tw, p);
        }
    }
}

// LDS slot of jj in the exchange buffer: jj + jj/8, one pad word after every 8.
// Unpadded, layouts C and D put 8 lanes of a wave on one bank (8-way conflicts);
// padded, no layout exceeds 2-way.  Each form is written out as a per-thread base
// plus a constant per element, so the padding costs no address arithmetic
// (a swizzle computed per element measured slower than the conflicts).
constexpr uint32_t kRowPadded = kRowC + kRowC / 8;

__device__ __forceinline__ uint32_t padded_index(int layout, uint32_t t, int e) {
    switch (layout) {
        case 0: return t + (t >> 3) + 288u * e;
        case 1: return 288u * (t >> 5) + (t & 31u) + ((t & 31u) >> 3) + 36u * e;
        case 2: return 36u * (t >> 2) + (t & 3u) + 4u * e + (e >> 1);
        default: return 9u * t + e;
    }
}

__device__ __forceinline__ void row_exchange(uint32_t* buf, uint32_t (&x)[8], int from, int to, uint32_t t) {
#pragma unroll
    for (int e = 0; e < 8; ++e) buf[padded_index(from, t, e)] = x[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 8; ++e) x[e] = buf[padded_index(to, t, e)];
    __syncthreads();
}

template <uint32_t N1>
__device__ __forceinline__ void rows_2048_body(uint32_t* __restrict__ planes, const uint32_t* __restrict__ bplanes,
                                               const PrimeConst* __restrict__ consts, const uint2* __restrict__ twf,
                                               const uint2* __restrict__ twi, uint32_t ps, uint32_t n2, uint32_t rows,
                                               uint32_t row, uint32_t k) {
    constexpr uint32_t n1 = N1;
    __shared__ uint32_t buf[2][kRowPadded];
    const uint32_t t = threadIdx.x;
    const uint32_t p = consts[k].p, pinv = consts[k].pinv;
    uint32_t* pa = planes + size_t(k) * ps + size_t(row) * kRowC;
    const uint32_t* pb = bplanes + size_t(k) * ps + size_t(row) * kRowC;
    const uint2* wf = twf + size_t(k) * n2;
    const uint2* wi = twi + size_t(k) * n2;
    // Table index of the group holding row-local jj at distance d:
    // (rows + row) * (C / 2d) + jj / 2d  ==  ((rows + row) * C + jj) >> log2(2d).
    const uint32_t base_index = (rows + row) * kRowC;
    uint32_t xa[8], xb[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        xa[e] = pa[t + 256 * e];
        xb[e] = pb[t + 256 * e];
    }
    row_forward_group(xa, 0, 3, t, 256, n1, base_index, wf, p);
    row_forward_group(xb, 0, 3, t, 256, n1, base_index, wf, p);
    row_exchange(buf[0], xa, 0, 1, t);
    row_exchange(buf[1], xb, 0, 1, t);
    row_forward_group(xa, 1, 3, t, 32, n1, base_index, wf, p);
    row_forward_group(xb, 1, 3, t, 32, n1, base_index, wf, p);
    row_exchange(buf[0], xa, 1, 2, t);
    row_exchange(buf[1], xb, 1, 2, t);
    row_forward_group(xa, 2, 3, t, 4, n1, base_index, wf, p);
    row_forward_group(xb, 2, 3, t, 4, n1, base_index, wf, p);
    row_exchange(buf[0], xa, 2, 3, t);
    row_exchange(buf[1], xb, 2, 3, t);
    row_forward_group(xa, 3, 2, t, 1, n1, base_index, wf, p);
    row_forward_group(xb, 3, 2, t, 1, n1, base_index, wf, p);

    // base case: thread t holds jj = 8t .. 8t+7, i.e. 8/n1 whole blocks.  Block g
    // is reduced mod (X^n1 - zeta_g): c_s = sum_{u<=s} a_u b_{s-u} + sum_{u>s}
    // a_u (zeta_g b_{s+n1-u}).  The wrapped factors zeta_g*b_v are formed once per
    // block (Shoup), so each output is one sum of n1 products, reduced four at a
    // time (a sum of four products of residues is below 4p^2 < 2p*2^32, in
    // redc_wide's range; 32-bit multiplies are the scarce resource here).
    uint32_t xc[8];
    constexpr int n1i = int(N1);
#pragma unroll
    for (int blk0 = 0; blk0 < 8; blk0 += n1i) {
        uint32_t bw[n1i];  // bw[v] = zeta_g * b_v for v >= 1
        bw[0] = 0;
        if constexpr (n1i > 1) {
            const uint32_t gb = (row * kRowC + 8 * t + blk0) / n1;  // global block index
            uint2 w{0, 0};
            if (n2 != 1) w = wf[n2 / 2 + gb / 2];  // zeta_{2g} = w, zeta_{2g+1} = -w
#pragma unroll
            for (int v = 1; v < n1i; ++v) {
                const uint32_t b = xb[blk0 + v];
                uint32_t z;
                if (n2 == 1) {
                    z = b ? p - b : 0;  // zeta = -1
                } else {
                    z = mul_shoup(b, w.x, w.y, p);
                    if (gb & 1) z = z ? p - z : 0;
                }
                bw[v] = z;
            }
        }
#pragma unroll
        for (int s1 = 0; s1 < n1i; ++s1) {
            uint64_t acc = 0;
            uint32_t value = 0;
#pragma unroll
            for (int u = 0; u < n1i; ++u) {
                const uint32_t f = u <= s1 ? xb[blk0 + s1 - u] : bw[s1 + n1i - u];
                acc += uint64_t(xa[blk0 + u]) * f;
                if ((u & 3) == 3 || u == n1i - 1) {
                    value = add_mod(value, redc_wide(acc, p, pinv), p);
                    acc = 0;
                }
            }
            xc[blk0 + s1] = value;
        }
    }

    row_inverse_group(xc, 3, 2, t, 1, n1, base_index, wi, p);
    row_exchange(buf[0], xc, 3, 2, t);
    row_inverse_group(xc, 2, 3, t, 4, n1, base_index, wi, p);
    row_exchange(buf[0], xc, 2, 1, t);
    row_inverse_group(xc, 1, 3, t, 32, n1, base_index, wi, p);
    row_exchange(buf[0], xc, 1, 0, t);
    row_inverse_group(xc, 0, 3, t, 256, n1, base_index, wi, p);
#pragma unroll
    for (int e = 0; e < 8; ++e) pa[t + 256 * e] = xc[e];
}

template <uint32_t N1>
__global__ __launch_bounds__(256) void ntt_rows_2048(uint32_t* __restrict__ planes,
                                                     const uint32_t* __restrict__ bplanes,
                                                     const PrimeConst* __restrict__ consts,
                                                     const uint2* __restrict__ twf, const uint2* __restrict__ twi,
                                                     uint32_t ps, uint32_t n2, uint32_t rows, uint32_t r) {
    rows_2048_body<N1>(planes, bplanes, consts, twf, twi, ps, n2, rows, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// 6. reconstruction: one lane per coefficient, for q of up to 62 limbs (T of at
// most 64 words), words in [j][n_pad] layout.  Output row k*256 + u is
// coefficient tile_at(ts, k)*256 + u.
// ---------------------------------------------------------------------------
struct ReconstructArgs {
    uint32_t n, n_pad, limbs, t_words;
    TileSet ts;
    uint32_t s_shift, k_shift;
    uint64_t mu;
};

__device__ __forceinline__ void reconstruct_body(const int64_t* __restrict__ words, uint32_t* __restrict__ tsc,
                                                 uint32_t* __restrict__ c, const uint32_t* __restrict__ q,
                                                 const uint32_t* __restrict__ neg_p,
                                                 const uint32_t* __restrict__ corr_t,
                                                 const uint32_t* __restrict__ corr_e, ReconstructArgs args,
                                                 uint32_t bx) {
    __shared__ uint32_t tile[64][65];
    // Output rows k0 .. k0+63 lie in one tile of the set: coefficients i0 .. i0+63.
    const uint32_t lane = threadIdx.x, k0 = bx * 64;
    const uint32_t i0 = tile_at(args.ts, k0 / 256) * 256 + k0 % 256, i = i0 + lane;
    const uint32_t ps = args.n_pad, L = args.limbs, tw = args.t_words;
    const bool valid = i < args.n;
    const uint32_t ic = valid ? i : 0;  // clamp reads of idle lanes

    // K = floor((E + 2^63) / 2^64), E = sum_j 2^32j words[L+1+j] + corr_e
    uint32_t ew[4];
    {
        int64_t carry = 0;
        for (uint32_t j = 0; j < 4; ++j) {
            int64_t x = (j < 3 ? words[size_t(L + 1 + j) * ps + ic] : 0) + int64_t(corr_e[j]) + carry;
            ew[j] = uint32_t(uint64_t(x));
            carry = x >> 32;
        }
    }
    const uint64_t kq = uint64_t(ew[2]) + (uint64_t(ew[3]) << 32) + ((uint64_t(ew[1]) + 0x80000000ull) >> 32);

    // Pass 1: T = sum_j 2^32j words[j] + corr_t + K*R, carried into 32-bit words.
    // One lane owns a whole coefficient, so memory-level parallelism comes from
    // running the loads ahead of the carry chain: two groups of kG1 words
    // alternate, one in flight while the other is carried.
    const uint32_t sw = args.s_shift / 32, sb = args.s_shift % 32;
    uint32_t top[3] = {0, 0, 0};
    {
        constexpr uint32_t kG1 = 32;
        int64_t da[kG1], db[kG1];
        int64_t carry = 0;
        auto fetch = [&](int64_t (&d)[kG1], uint32_t j0) {
#pragma unroll
            for (uint32_t u = 0; u < kG1; ++u) {
                const uint32_t j = j0 + u;
                d[u] = j <= L && j < tw ? words[size_t(j) * ps + ic] : 0;
            }
        };
        auto carry_group = [&](const int64_t (&d)[kG1], uint32_t j0) {
#pragma unroll
            for (uint32_t u = 0; u < kG1; ++u) {
                const uint32_t j = j0 + u;
                if (j >= tw) break;
                int64_t x = d[u] + int64_t(corr_t[j]) + carry;
                if (j < L) x += int64_t(kq) * int64_t(neg_p[j]);
                const uint32_t tj = uint32_t(uint64_t(x));
                carry = x >> 32;
                if (valid) tsc[size_t(j) * ps + ic] = tj;
                // the three words Barrett reads; compares, not a runtime index (which would spill)
                if (j == sw) top[0] = tj;
                if (j == sw + 1) top[1] = tj;
                if (j == sw + 2) top[2] = tj;
            }
        };
        fetch(da, 0);
        for (uint32_t j0 = 0; j0 < tw; j0 += 2 * kG1) {
            if (j0 + kG1 < tw) fetch(db, j0 + kG1);
            carry_group(da, j0);
            if (j0 + kG1 >= tw) break;
            if (j0 + 2 * kG1 < tw) fetch(da, j0 + 2 * kG1);
            carry_group(db, j0 + kG1);
        }
    }
    const unsigned __int128 topv = (unsigned __int128)top[0] | ((unsigned __int128)top[1] << 32) |
                                   ((unsigned __int128)top[2] << 64);
    const uint64_t t_scaled = uint64_t(topv >> sb);
    const uint64_t qe = uint64_t(((unsigned __int128)t_scaled * args.mu) >> args.k_shift);

    // Pass 2: rem = T - qe*q (in [0, 2q)) into an LDS row per lane, in two
    // halves of 32 words (the second in flight while the first is reduced); a
    // second borrow chain tracks rem - q, and when rem >= q the row is reduced
    // once more in LDS.  The block's rows are then stored as one contiguous run
    // of words: c is written once and never read back, so it may be pinned host
    // memory.
    {
        uint64_t pc = 0;
        uint32_t bo = 0, bo2 = 0;
        constexpr uint32_t kH = 32;
        uint32_t ta[kH], tb[kH];
        auto fetch = [&](uint32_t (&t)[kH], uint32_t j0) {
#pragma unroll
            for (uint32_t u = 0; u < kH; ++u) t[u] = j0 + u < tw ? tsc[size_t(j0 + u) * ps + ic] : 0;
        };
        auto reduce_half = [&](const uint32_t (&t)[kH], uint32_t j0, uint32_t col0) {
#pragma unroll
            for (uint32_t u = 0; u < kH; ++u) {
                const uint32_t j = j0 + u;
                if (j >= tw) break;
                const uint32_t qj = j < L ? q[j] : 0;
                const unsigned __int128 pr = (unsigned __int128)qe * qj + pc;
                const uint32_t x = uint32_t(pr);
                pc = uint64_t(pr >> 32);
                const uint64_t d = uint64_t(t[u]) - x - bo;
                const uint32_t rj = uint32_t(d);
                bo = uint32_t(d >> 63);
                const uint64_t d2 = uint64_t(rj) - qj - bo2;
                bo2 = uint32_t(d2 >> 63);
                tile[lane][col0 + u] = rj;
            }
        };
        fetch(ta, 0);
        if (kH < tw) fetch(tb, kH);
        reduce_half(ta, 0, 0);
        if (kH < tw) reduce_half(tb, kH, kH);
        if (bo2 == 0) {
            uint32_t borrow = 0;
            for (uint32_t j = 0; j < L; ++j) {
                const uint64_t d = uint64_t(tile[lane][j]) - q[j] - borrow;
                tile[lane][j] = uint32_t(d);
                borrow = uint32_t(d >> 63);
            }
        }
        __syncthreads();
        const uint32_t rows = i0 < args.n ? (args.n - i0 < 64 ? args.n - i0 : 64) : 0;
        uint32_t* out = c + size_t(k0) * L;
        const uint32_t dq = 64 / L, dr = 64 % L;  // (row, col) of word k, stepped by 64 words
        uint32_t row = lane / L, col = lane % L;
        for (uint32_t k = lane; k < rows * L; k += 64) {
            out[k] = tile[row][col];
            row += dq;
            col += dr;
            if (col >= L) {
                col -= L;
                ++row;
            }
        }
    }
}

__global__ __launch_bounds__(64) void reconstruct(const int64_t* __restrict__ words,
                                                  uint32_t* __restrict__ tsc, uint32_t* __restrict__ c,
                                                  const uint32_t* __restrict__ q,
                                                  const uint32_t* __restrict__ neg_p,
                                                  const uint32_t* __restrict__ corr_t,
                                                  const uint32_t* __restrict__ corr_e,
                                                  ReconstructArgs args) {
    reconstruct_body(words, tsc, c, q, neg_p, corr_t, corr_e, args, blockIdx.x);
}

// ---------------------------------------------------------------------------
// 6'. reconstruction, one wave per coefficient (q wider than 62 limbs): the
// coefficient's CRT word sums are coefficient-major (words[i][0 .. L+3]) and
// each lane takes one of 64 consecutive words per round, so every round is one
// 512-byte load and one 256-byte store, and N waves (not N lanes) keep the
// memory system busy.
//
//   K  = floor((E + 2^63) / 2^64), E = sum_j 2^32j D_{L+1+j} + corr_e
//   x_j = D_j + corr_t_j + K*negp_j - (Qe*q)_j            (|x_j| < 2^56)
//   R  = sum_j x_j 2^32j = T - Qe*q
// Qe is the Barrett quotient of T estimated from the four words of x from
// m = floor(s/32) - 1 up (all words from 0 when s < 64): the words below m
// shift that estimate by under 2^25 * 2^32m, i.e. under 2^-7 at the 2^s
// scale, so Ts is off by at most one and Qe lies in [Q-2, Q+1].  One pass
// normalizes x into 32-bit words (carry out of each round passed to the next)
// and compares R with q on the way; in the rare case R < 0 or R >= q, Qe is
// moved by one and the pass repeats, rewriting c.
// ---------------------------------------------------------------------------
struct WaveReconstructArgs {
    uint32_t n, ws, limbs, t_words;
    TileSet ts;  // output row k*256 + u is coefficient tile_at(ts, k)*256 + u
    uint32_t top_m, top_e, k_shift;
    uint64_t mu;
};

__device__ __forceinline__ int32_t from_lane_below(int32_t v, uint32_t lane) {
    const int32_t u = __shfl_up(v, 1, 64);
    return lane ? u : 0;
}

__device__ __forceinline__ int64_t read_lane64(int64_t v, int l) {
    const uint32_t lo = __builtin_amdgcn_readlane(int(uint32_t(v)), l);
    const uint32_t hi = __builtin_amdgcn_readlane(int(uint32_t(uint64_t(v) >> 32)), l);
    return int64_t((uint64_t(hi) << 32) | lo);
}

// Normalizes one round of 64 signed words x (lane j: |x_j| < 2^57) plus the
// wave-uniform carry into word 0 into 32-bit words; returns lane j's word and
// leaves in `carry` the carry into the next round's word 0.
__device__ __forceinline__ uint32_t normalize_round(int64_t x, int64_t& carry, uint32_t lane) {
    if (lane == 0) x += carry;
    const int32_t hi = int32_t(x >> 32);                              // |hi| < 2^25
    const int64_t v = int64_t(uint32_t(x)) + from_lane_below(hi, lane);  // (-2^25, 2^32 + 2^25)
    const int32_t cj = int32_t(v >> 32);                              // -1, 0 or 1
    const int64_t u = int64_t(uint32_t(v)) + from_lane_below(cj, lane);  // [-1, 2^32]
    const int32_t hi63 = __builtin_amdgcn_readlane(hi, 63), c63 = __builtin_amdgcn_readlane(cj, 63);
    uint32_t w = uint32_t(u);
    int64_t ripple = 0;
    if (__ballot(u < 0 || u > int64_t(0xFFFFFFFFu))) {
        // a word of -1 or 2^32 (rare): ripple its carry up the round
        int64_t r = 0;
        for (int l = 0; l < 64; ++l) {
            const int64_t t = read_lane64(u, l) + r;
            if (lane == uint32_t(l)) w = uint32_t(t);
            r = t >> 32;
        }
        ripple = r;
    }
    carry = int64_t(hi63) + c63 + ripple;
    return w;
}

__device__ __forceinline__ void reconstruct_waves_body(const int64_t* __restrict__ words, uint32_t* __restrict__ c,
                                                       const uint32_t* __restrict__ q,
                                                       const uint32_t* __restrict__ neg_p,
                                                       const uint32_t* __restrict__ corr_t,
                                                       const uint32_t* __restrict__ corr_e,
                                                       WaveReconstructArgs args, uint32_t bx) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t k = bx * 4 + (threadIdx.x >> 6);  // output row
    if (k >= args.ts.count * 256) return;  // whole waves
    const uint32_t i = tile_at(args.ts, k / 256) * 256 + k % 256;
    if (i >= args.n) return;
    const uint32_t L = args.limbs, tw = args.t_words;
    const int64_t* d = words + size_t(i) * args.ws;
    uint32_t* out = c + size_t(k) * L;

    // K = floor((E + 2^63) / 2^64)
    uint32_t ew[4];
    {
        int64_t carry = 0;
        for (uint32_t j = 0; j < 4; ++j) {
            const int64_t x = (j < 3 ? d[L + 1 + j] : 0) + int64_t(corr_e[j]) + carry;
            ew[j] = uint32_t(uint64_t(x));
            carry = x >> 32;
        }
    }
    const int64_t kq = int64_t(uint64_t(ew[2]) + (uint64_t(ew[3]) << 32) + ((uint64_t(ew[1]) + 0x80000000ull) >> 32));

    // Barrett estimate from the top four words: S = sum_t x_{m+t} 2^32t mod 2^128
    // equals (T - [words below m]) / 2^32m exactly (|S| < 2^126).
    uint64_t qe;
    {
        unsigned __int128 sum = 0;
        for (uint32_t t = 0; t < 4; ++t) {
            const uint32_t j = args.top_m + t;
            if (j >= tw) break;
            const int64_t x = (j <= L ? d[j] : 0) + int64_t(corr_t[j]) + (j < L ? kq * int64_t(neg_p[j]) : 0);
            sum += (unsigned __int128)(__int128)x << (32 * t);
        }
        __int128 ts = __int128(sum) >> args.top_e;
        if (ts < 0) ts = 0;
        qe = uint64_t(((unsigned __int128)uint64_t(ts) * args.mu) >> args.k_shift);
    }

    // A round's inputs (lane j of words j0 .. j0+63): D_j + corr_t_j + K*negp_j and
    // the three words of q that word j of Qe*q needs.  Every load is
    // unconditional, its index clamped into its table and the value selected
    // afterwards, so a round carries no divergent branches, and the loads of a
    // round are issued two rounds before it is normalized.  K < 2^32 (it is
    // below the number of primes), so K*negp_j is one 32 x 32-bit product.
    struct Round {  // the raw loads; masked and combined only when the round is used
        int64_t d;
        uint32_t ct, np, qj, qm1, qm2;
    };
    const uint32_t k32 = uint32_t(kq);
    auto fetch = [&](uint32_t j0) {
        const uint32_t j = j0 + lane;
        const uint32_t jq = j < L ? j : L - 1, jq1 = j - 1 < L ? j - 1 : 0, jq2 = j - 2 < L ? j - 2 : 0;
        return Round{d[j <= L ? j : L], corr_t[j < tw ? j : tw - 1], neg_p[jq], q[jq], q[jq1], q[jq2]};
    };
    for (int pass = 0; pass < 6; ++pass) {
        const uint64_t q0 = uint32_t(qe), q1 = uint32_t(qe >> 32);
        int64_t carry = 0;
        int cmp = 0;  // sign of R - q over the words so far (decided by the highest differing word)
        auto round = [&](const Round& in, uint32_t j0) {
            const uint32_t j = j0 + lane;
            const uint32_t qj = j < L ? in.qj : 0;
            const uint32_t qm1 = j >= 1 && j - 1 < L ? in.qm1 : 0;
            const uint32_t qm2 = j >= 2 && j - 2 < L ? in.qm2 : 0;
            // word j of Qe*q: lo(q0 q_j) + hi(q0 q_{j-1}) + lo(q1 q_{j-1}) + hi(q1 q_{j-2})
            const uint64_t pa = q0 * qj, pb = q0 * qm1, pc = q1 * qm1, pd = q1 * qm2;
            const int64_t sub = int64_t(uint32_t(pa)) + int64_t(pb >> 32) + int64_t(uint32_t(pc)) + int64_t(pd >> 32);
            const int64_t kr = j < L ? int64_t(uint64_t(k32) * in.np) : 0;
            const int64_t x = j < tw ? (j <= L ? in.d : 0) + int64_t(in.ct) + kr - sub : 0;
            const uint32_t w = normalize_round(x, carry, lane);
            if (j < L) out[j] = w;
            const uint64_t differ = __ballot(w != qj);
            if (differ) {
                const int top = 63 - __clzll(differ);
                cmp = uint32_t(__builtin_amdgcn_readlane(int(w), top)) > uint32_t(__builtin_amdgcn_readlane(int(qj), top))
                          ? 1
                          : -1;
            }
        };
        // Three register sets in rotation (no copies between them, which would
        // wait for the loads): each set is refilled right after its round.
        Round r0 = fetch(0), r1 = fetch(64), r2;
        for (uint32_t j0 = 0; j0 < tw; j0 += 192) {
            r2 = fetch(j0 + 128);
            round(r0, j0);
            if (j0 + 64 >= tw) break;
            r0 = fetch(j0 + 192);
            round(r1, j0 + 64);
            if (j0 + 128 >= tw) break;
            r1 = fetch(j0 + 256);
            round(r2, j0 + 128);
        }
        if (carry < 0) {
            --qe;  // R < 0
        } else if (carry > 0 || cmp >= 0) {
            ++qe;  // R >= q
        } else {
            break;
        }
    }
}

__global__ __launch_bounds__(256) void reconstruct_waves(const int64_t* __restrict__ words, uint32_t* __restrict__ c,
                                                         const uint32_t* __restrict__ q,
                                                         const uint32_t* __restrict__ neg_p,
                                                         const uint32_t* __restrict__ corr_t,
                                                         const uint32_t* __restrict__ corr_e,
                                                         WaveReconstructArgs args) {
    reconstruct_waves_body(words, c, q, neg_p, corr_t, corr_e, args, blockIdx.x);
}

// ---------------------------------------------------------------------------
// 6''. Grouped launches: one stage of many products in one launch.  Entry j of
// a launch owns blocks [start[j], start[j+1]) and runs them as the blocks of
// product item[j]'s own launch of that stage (kernels.hpp, GroupItem).
// ---------------------------------------------------------------------------
struct GroupLaunch {
    const GroupItem* items;
    const uint32_t* start;  // count + 1 block offsets
    const uint32_t* item;   // the product of each entry
    uint32_t count;
};

// The product owning this block (a binary search over the offsets, the same
// for every thread of the block) and the block's index in that product's grid.
__device__ __forceinline__ const GroupItem& group_entry(const GroupLaunch& g, uint32_t& local) {
    const uint32_t b = blockIdx.x;
    uint32_t lo = 0, hi = g.count;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (g.start[mid] <= b)
            lo = mid;
        else
            hi = mid;
    }
    local = b - g.start[lo];
    return g.items[g.item[lo]];
}

__device__ __forceinline__ TileSet every_tile(const GroupItem& it) { return tile_range(0, it.n_pad / kTile); }

__global__ __launch_bounds__(256) void pack_group(GroupLaunch g) {
    uint32_t local;
    const GroupItem& it = group_entry(g, local);
    pack_body(it.a, it.b, reinterpret_cast<uint32_t*>(it.packed), it.n, it.limbs, it.conv_k / kTK, it.n_pad / kTile,
              every_tile(it), 0, 2, local);
}

__global__ __launch_bounds__(512) void conversion_group(GroupLaunch g) {
    uint32_t local;
    const GroupItem& it = group_entry(g, local);
    conversion_body(it.wconv, it.packed, it.consts, it.planes, it.n, it.n_pad, it.r, it.r_pad, it.conv_rows,
                    it.conv_k / kTK, every_tile(it), 0, 2, local);
}

// grid per product: (row_width / 256, r, 2)
template <int ROWS>
__global__ __launch_bounds__(256) void forward_reg_group(GroupLaunch g) {
    uint32_t local;
    const GroupItem& it = group_entry(g, local);
    const uint32_t nx = it.row_width / 256, bx = local % nx, rest = local / nx;
    forward_reg_body<ROWS>(it.planes, it.consts, it.tw_fwd, it.n_pad, it.n2, it.row_width, it.r_pad, 0, 0, bx,
                           rest % it.r, rest / it.r);
}

// grid per product: (rows, r)
template <uint32_t N1>
__global__ __launch_bounds__(256) void rows_2048_group(GroupLaunch g) {
    uint32_t local;
    const GroupItem& it = group_entry(g, local);
    rows_2048_body<N1>(it.planes, it.planes + size_t(it.r_pad) * it.n_pad, it.consts, it.tw_fwd, it.tw_inv, it.n_pad,
                       it.n2, it.rows, local % it.rows, local / it.rows);
}

__global__ __launch_bounds__(256) void rows_product_group(GroupLaunch g) {
    extern __shared__ uint32_t lds32[];
    uint32_t local;
    const GroupItem& it = group_entry(g, local);
    rows_product_body(it.planes, it.planes + size_t(it.r_pad) * it.n_pad, it.consts, it.tw_fwd, it.tw_inv, it.n_pad,
                      it.n1, it.n2, it.row_width, it.rows, lds32, local % it.rows, local / it.rows);
}

// grid per product: (row_width / 16, r_pad / 16)
template <int ROWS>
__global__ __launch_bounds__(256) void inverse_tiles_group(GroupLaunch g) {
    uint32_t local;
    const GroupItem& it = group_entry(g, local);
    const uint32_t nx = it.row_width / 16;
    inverse_tiles_body<ROWS>(it.planes, it.consts, it.tw_inv, it.n_pad, it.n, it.n2, it.row_width, it.r,
                             it.r_pad / 16, it.rtiles, 0, local % nx, local / nx, nx);
}

template <bool kByCoefficient>
__global__ __launch_bounds__(512) void crt_group(GroupLaunch g) {
    uint32_t local;
    const GroupItem& it = group_entry(g, local);
    crt_body<kByCoefficient>(it.wcrt, it.rtiles, it.words, it.n, it.n_pad, it.crt_rows, it.crt_k / kTK, it.crt_words,
                             it.ws, every_tile(it), local);
}

__global__ __launch_bounds__(64) void reconstruct_group(GroupLaunch g) {
    uint32_t local;
    const GroupItem& it = group_entry(g, local);
    const ReconstructArgs args{it.n, it.n_pad, it.limbs, it.t_words, every_tile(it), it.s_shift, it.k_shift, it.mu};
    reconstruct_body(it.words, it.t_scratch, it.c, it.q, it.neg_p, it.corr_t, it.corr_e, args, local);
}

__global__ __launch_bounds__(256) void reconstruct_waves_group(GroupLaunch g) {
    uint32_t local;
    const GroupItem& it = group_entry(g, local);
    const WaveReconstructArgs args{it.n, it.ws, it.limbs, it.t_words, every_tile(it),
                                   it.top_m, it.top_e, it.k_shift, it.mu};
    reconstruct_waves_body(it.words, it.c, it.q, it.neg_p, it.corr_t, it.corr_e, args, local);
}

__global__ void fill_zero(uint32_t* __restrict__ c, size_t words) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < words) c[i] = 0;
}

// ---------------------------------------------------------------------------
// 7. plan-time tables (host equivalents in plan.cpp, build_plan).
// ---------------------------------------------------------------------------
// x*y mod p for x, y < p < 2^31 with rcp = floor((2^64-1)/p): for t = xy < 2^62
// the estimate hi64(t*rcp) lies in (t/p - 2, t/p], so two corrections finish.
__device__ __forceinline__ uint32_t mulmod_rcp(uint32_t x, uint32_t y, uint32_t p, uint64_t rcp) {
    const uint64_t t = uint64_t(x) * y;
    uint64_t r = t - __umul64hi(t, rcp) * p;
    if (r >= p) r -= p;
    if (r >= p) r -= p;
    return uint32_t(r);
}

__device__ __forceinline__ uint32_t powmod_rcp(uint32_t b, uint32_t e, uint32_t p, uint64_t rcp) {
    uint32_t r = 1;
    for (; e; e >>= 1) {
        if (e & 1) r = mulmod_rcp(r, b, p, rcp);
        b = mulmod_rcp(b, b, p, rcp);
    }
    return r;
}

// Shoup quotient floor(w * 2^32 / p) for w < p, by the same estimate.
__device__ __forceinline__ uint32_t shoup_rcp(uint32_t w, uint32_t p, uint64_t rcp) {
    const uint64_t t = uint64_t(w) << 32;
    uint64_t q = __umul64hi(t, rcp), r = t - q * p;
    if (r >= p) { r -= p; ++q; }
    if (r >= p) ++q;
    return uint32_t(q);
}

// beta_k = n2^-1 * 2^32 * P_k^-1 mod p_k, with P_k mod p_k = prod_{j != k} p_j
// formed by one block per prime (strided products, then a tree).
__global__ __launch_bounds__(256) void table_beta(const PrimeConst* __restrict__ consts,
                                                  const uint64_t* __restrict__ rcp, uint32_t* __restrict__ beta,
                                                  uint32_t r, uint32_t n2) {
    __shared__ uint32_t part[256];
    const uint32_t k = blockIdx.x, t = threadIdx.x;
    const uint32_t p = consts[k].p;
    const uint64_t rc = rcp[k];
    uint32_t prod = 1;
    for (uint32_t j = t; j < r; j += 256)
        if (j != k) prod = mulmod_rcp(prod, consts[j].p % p, p, rc);
    part[t] = prod;
    __syncthreads();
    for (uint32_t w = 128; w; w >>= 1) {
        if (t < w) part[t] = mulmod_rcp(part[t], part[t + w], p, rc);
        __syncthreads();
    }
    if (t == 0) {
        const uint32_t radix = uint32_t((uint64_t(1) << 32) % p);
        const uint32_t n2_inv = powmod_rcp(n2 % p, p - 2, p, rc);
        const uint32_t pk_inv = powmod_rcp(part[0], p - 2, p, rc);
        beta[k] = mulmod_rcp(mulmod_rcp(n2_inv, radix, p, rc), pk_inv, p, rc);
    }
}

// tw[k][j] = (psi^e, Shoup quotient) with e = bitrev(j) over log2(n2) bits; the
// inverse table uses psi^-1.  grid (column blocks, r).
__global__ __launch_bounds__(256) void table_twiddles(const PrimeConst* __restrict__ consts,
                                                      const uint64_t* __restrict__ rcp,
                                                      const uint32_t* __restrict__ psi, uint32_t* __restrict__ fwd,
                                                      uint32_t* __restrict__ inv, uint32_t n2, uint32_t log_n2) {
    __shared__ uint32_t pf[32], pi[32];  // psi^(2^i), psi^-(2^i)
    const uint32_t k = blockIdx.y;
    const uint32_t p = consts[k].p;
    const uint64_t rc = rcp[k];
    if (threadIdx.x == 0) {
        uint32_t a = psi[2 * k], b = psi[2 * k + 1];
        for (uint32_t i = 0; i < log_n2; ++i) {
            pf[i] = a;
            pi[i] = b;
            a = mulmod_rcp(a, a, p, rc);
            b = mulmod_rcp(b, b, p, rc);
        }
    }
    __syncthreads();
    for (uint32_t j = blockIdx.x * 256 + threadIdx.x; j < n2; j += gridDim.x * 256) {
        const uint32_t e = log_n2 ? __brev(j) >> (32 - log_n2) : 0;
        uint32_t w = 1, wi = 1;
        for (uint32_t i = 0; i < log_n2; ++i)
            if ((e >> i) & 1) {
                w = mulmod_rcp(w, pf[i], p, rc);
                wi = mulmod_rcp(wi, pi[i], p, rc);
            }
        const size_t at = (size_t(k) * n2 + j) * 2;
        *reinterpret_cast<uint2*>(fwd + at) = make_uint2(w, shoup_rcp(w, p, rc));
        *reinterpret_cast<uint2*>(inv + at) = make_uint2(wi, shoup_rcp(wi, p, rc));
    }
}

// Conversion weights of operand op for prime k: columns j < 4L hold
// lambda * 256^j mod p (lambda = 2^32 for a, 2^32 * beta_k for b), centered and
// split into four signed digits (rows 4k+d); columns j >= 4L are zero.  The
// block also forms off = 128 * sum_j centered_j mod p.  Each thread writes four
// consecutive columns of a row as one word into the tiled matrix (pre-zeroed).
// grid (r, 2).
__global__ __launch_bounds__(256) void table_conversion(PrimeConst* __restrict__ consts,
                                                        const uint64_t* __restrict__ rcp,
                                                        const uint32_t* __restrict__ beta,
                                                        int8_t* __restrict__ wtiles, uint32_t conv_rows,
                                                        uint32_t conv_k, uint32_t limbs) {
    __shared__ int64_t part[256];
    const uint32_t k = blockIdx.x, op = blockIdx.y, t = threadIdx.x;
    const uint32_t p = consts[k].p;
    const uint64_t rc = rcp[k];
    const uint32_t radix = uint32_t((uint64_t(1) << 32) % p);
    const uint32_t lambda = op ? mulmod_rcp(radix, beta[k], p, rc) : radix;
    const uint32_t ksteps = conv_k / kTK, num_m = conv_rows / kTile, kbytes = 4 * limbs;
    const uint32_t step = powmod_rcp(256, 4 * 256, p, rc);  // 256^(columns per block pass)
    uint32_t cur = mulmod_rcp(lambda, powmod_rcp(256, 4 * t, p, rc), p, rc);
    int64_t sum = 0;
    for (uint32_t j0 = 4 * t; j0 < conv_k; j0 += 4 * 256) {
        uint32_t word[4] = {0, 0, 0, 0};
        uint32_t w = cur;
        for (uint32_t i = 0; i < 4; ++i) {
            if (j0 + i < kbytes) {
                int64_t x = w > p / 2 ? int64_t(w) - p : int64_t(w);
                sum += x;
                for (uint32_t d = 0; d < 4; ++d) {
                    const int64_t e = ((x + 128) & 255) - 128;
                    word[d] |= uint32_t(uint8_t(int8_t(e))) << (8 * i);
                    x = (x - e) >> 8;
                }
            }
            w = mulmod_rcp(w, 256, p, rc);
        }
        for (uint32_t d = 0; d < 4; ++d) {
            if (!word[d]) continue;
            const uint32_t row = 4 * k + d;
            const size_t at = ((size_t(op * num_m + row / kTile) * ksteps + j0 / kTK) * kTile + row % kTile) * kTK +
                              j0 % kTK;
            *reinterpret_cast<uint32_t*>(wtiles + at) = word[d];
        }
        cur = mulmod_rcp(cur, step, p, rc);
    }
    part[t] = sum;
    __syncthreads();
    for (uint32_t w = 128; w; w >>= 1) {
        if (t < w) part[t] += part[t + w];
        __syncthreads();
    }
    if (t == 0) {
        int64_t off = (128 * part[0]) % int64_t(p);
        if (off < 0) off += p;
        if (op) consts[k].off_b = uint32_t(off);
        else consts[k].off_a = uint32_t(off);
    }
}

// CRT weights of prime k (one thread each): column 4k+d of row c holds signed
// digit c-d of M_k (rows 0 .. 4L+3) and, from row 4(L+1) on, signed digit
// c-4(L+1)-d of rcp_k.  Row c's four columns are one word whose byte d is digit
// c-d: a sliding window over the digit stream.  Matrix pre-zeroed.
__global__ __launch_bounds__(64) void table_crt(const uint32_t* __restrict__ m_mod_q,
                                                const uint64_t* __restrict__ rcp, int8_t* __restrict__ wtiles,
                                                uint32_t r, uint32_t limbs, uint32_t crt_k) {
    const uint32_t k = blockIdx.x * 64 + threadIdx.x;
    if (k >= r) return;
    const uint32_t ksteps = crt_k / kTK, col = 4 * k;
    auto put = [&](uint32_t row, uint32_t word) {
        if (!word) return;
        const size_t at = ((size_t(row / kTile) * ksteps + col / kTK) * kTile + row % kTile) * kTK + col % kTK;
        *reinterpret_cast<uint32_t*>(wtiles + at) = word;
    };
    const uint32_t* m = m_mod_q + size_t(k) * limbs;
    const uint32_t md = 4 * limbs + 1;
    uint32_t window = 0;
    int carry = 0;
    for (uint32_t c = 0; c < md + 3; ++c) {
        int digit = 0;
        if (c < md) {
            const int x = (c / 4 < limbs ? int((m[c / 4] >> (8 * (c % 4))) & 255u) : 0) + carry;
            carry = x >= 128;
            digit = carry ? x - 256 : x;
        }
        window = (window << 8) | uint8_t(int8_t(digit));
        put(c, window);
    }
    const uint64_t rv = rcp[k];
    const uint32_t base = 4 * (limbs + 1);
    window = 0;
    carry = 0;
    for (uint32_t c = 0; c < 6 + 3; ++c) {
        int digit = 0;
        if (c < 6) {
            const int x = int((rv >> (8 * c)) & 255u) + carry;
            carry = x >= 128;
            digit = carry ? x - 256 : x;
        }
        window = (window << 8) | uint8_t(int8_t(digit));
        put(base + c, window);
    }
}

uint32_t column_width(const Shape& s) {
    uint32_t cw = 8192 / s.rows;
    if (cw > s.row_width) cw = s.row_width;
    if (cw < 1) cw = 1;
    return cw;
}

// Every tile of `ts` must be one of the plan's: a set past them would have the
// kernels read and write past the working buffers.
void check_tiles(const Shape& s, const TileSet& ts, const char* what) {
    const uint64_t ntiles = s.n_pad / kTile;
    const bool ok = ts.count != 0 && ts.w != 0 && ts.count % ts.w == 0 &&
                    uint64_t(ts.t0) + uint64_t(ts.count / ts.w - 1) * ts.stride + (ts.w - 1) < ntiles;
    if (!ok) throw std::invalid_argument(std::string(what) + ": coefficient tiles outside the plan's");
}

// Strips [strip_begin, strip_end) of 256 columns; (0, ~0u) is every column.
bool whole_columns(const Shape& s, uint32_t strip_begin, uint32_t strip_end, const char* what) {
    if (strip_begin == 0 && strip_end == ~0u) return true;
    if (s.row_width % 256 != 0 || s.rows < 2 || s.rows > 64 || strip_begin >= strip_end ||
        strip_end > s.row_width / 256)
        throw std::invalid_argument(std::string(what) + ": column strips outside the transform's columns");
    return false;
}

}  // namespace

size_t packed_operand_bytes(const Shape& s) { return size_t(2) * s.n_pad * s.conv_k; }

void launch_pack(const Shape& s, const uint32_t* a, const uint32_t* b, int8_t* packed, const TileSet& ts,
                 uint32_t op0, uint32_t nops, hipStream_t stream) {
    check_tiles(s, ts, "pack_operands");
    const uint32_t ksteps = s.conv_k / kTK, ntiles = s.n_pad / kTile;
    const size_t rows = size_t(nops) * ts.count * kTile;  // 4 waves per block, 1, 2 or 4 rows per wave
    const uint32_t per_block = 4 * pack_rows_per_wave(ksteps);
    dim3 grid(uint32_t((rows + per_block - 1) / per_block));
    hipLaunchKernelGGL(pack_operands, grid, dim3(256), 0, stream, a, b, reinterpret_cast<uint32_t*>(packed), s.n,
                       s.limbs, ksteps, ntiles, ts, op0, nops);
    check_launch("pack_operands");
}

void launch_conversion(const Shape& s, const DeviceTables& t, const int8_t* packed, uint32_t* planes,
                       const TileSet& ts, uint32_t op0, uint32_t nops, hipStream_t stream) {
    check_tiles(s, ts, "conversion_gemm");
    const uint32_t blocks = (s.conv_rows / kTile) * nops * ts.count;
    hipLaunchKernelGGL(conversion_gemm, dim3(blocks), dim3(kGemmThreads), 0, stream, t.wconv, packed,
                       t.consts, planes, s.n, s.n_pad, s.r, s.r_pad, s.conv_rows, s.conv_k / kTK, ts, op0, nops);
    check_launch("conversion_gemm");
}

void launch_ntt_forward_columns(const Shape& s, const DeviceTables& t, uint32_t* planes, uint32_t op0,
                                uint32_t nops, hipStream_t stream, uint32_t strip_begin, uint32_t strip_end) {
    const bool whole = whole_columns(s, strip_begin, strip_end, "ntt_columns_forward");
    if (s.rows < 2) return;
    const uint32_t ps = s.n_pad, r = s.r_pad, C = s.row_width;
    if (C % 256 == 0 && s.rows <= 64) {
        const uint32_t s0 = whole ? 0 : strip_begin, s1 = whole ? C / 256 : strip_end;
        dim3 grid(s1 - s0, s.r, nops);
        switch (s.rows) {
#define PM_FWD(R)                                                                                                \
    case R:                                                                                                      \
        hipLaunchKernelGGL(ntt_columns_forward_reg<R>, grid, dim3(256), 0, stream, planes, t.consts, t.tw_fwd, ps, \
                           s.n2, C, r, op0, s0 * 256);                                                           \
        break;
            PM_FWD(2) PM_FWD(4) PM_FWD(8) PM_FWD(16) PM_FWD(32) PM_FWD(64)
#undef PM_FWD
        }
    } else {
        const uint32_t cw = column_width(s);
        dim3 grid(C / cw, s.r, nops);
        hipLaunchKernelGGL(ntt_columns_forward, grid, dim3(256), size_t(s.rows) * cw * 4, stream, planes, t.consts,
                           t.tw_fwd, ps, s.n2, C, s.rows, cw, r, op0);
    }
    check_launch("ntt_columns_forward");
}

void launch_ntt_rows_product(const Shape& s, const DeviceTables& t, uint32_t* planes, const uint32_t* bplanes,
                             hipStream_t stream) {
    const uint32_t ps = s.n_pad, r = s.r_pad;
    dim3 grid(s.rows, s.r);
    if (s.row_width == kRowC && s.n1 <= 8) {
        switch (s.n1) {
#define PM_ROWS(B)                                                                                              \
    case B:                                                                                                     \
        hipLaunchKernelGGL(ntt_rows_2048<B>, grid, dim3(256), 0, stream, planes, bplanes, t.consts, t.tw_fwd,     \
                           t.tw_inv, ps, s.n2, s.rows, r);                                                      \
        break;
            PM_ROWS(1) PM_ROWS(2) PM_ROWS(4) PM_ROWS(8)
#undef PM_ROWS
        }
    } else {
        hipLaunchKernelGGL(ntt_rows_product, grid, dim3(256), size_t(2) * s.row_width * 4, stream, planes, bplanes,
                           t.consts, t.tw_fwd, t.tw_inv, ps, s.n1, s.n2, s.row_width, s.rows, r);
    }
    check_launch("ntt_rows_product");
}

void launch_ntt_inverse_columns(const Shape& s, const DeviceTables& t, uint32_t* planes, int8_t* rtiles,
                                hipStream_t stream, uint32_t strip_begin, uint32_t strip_end) {
    const bool whole = whole_columns(s, strip_begin, strip_end, "ntt_columns_inverse");
    const uint32_t ps = s.n_pad, ksteps = s.r_pad / 16;
    if (s.row_width % 16 == 0 && s.rows <= 64) {
        const uint32_t g0 = whole ? 0 : strip_begin * 16, g1 = whole ? s.row_width / 16 : strip_end * 16;
        dim3 grid(g1 - g0, s.r_pad / 16);
        switch (s.rows) {
#define PM_INV(R)                                                                                              \
    case R:                                                                                                    \
        hipLaunchKernelGGL(ntt_columns_inverse_tiles<R>, grid, dim3(256), 0, stream, planes, t.consts, t.tw_inv, \
                           ps, s.n, s.n2, s.row_width, s.r, ksteps, rtiles, g0);                               \
        break;
            PM_INV(1) PM_INV(2) PM_INV(4) PM_INV(8) PM_INV(16) PM_INV(32) PM_INV(64)
#undef PM_INV
        }
    } else {
        // large or narrow shapes: LDS column stages, then a pure transpose into the tiles
        if (s.rows >= 2) {
            const uint32_t cw = column_width(s);
            dim3 grid(s.row_width / cw, s.r);
            hipLaunchKernelGGL(ntt_columns_inverse, grid, dim3(256), size_t(s.rows) * cw * 4, stream, planes,
                               t.consts, t.tw_inv, ps, s.n2, s.row_width, s.rows, cw);
            check_launch("ntt_columns_inverse");
        }
        dim3 grid((s.n + 15) / 16, s.r_pad / 16);
        hipLaunchKernelGGL(ntt_columns_inverse_tiles<1>, grid, dim3(256), 0, stream, planes, t.consts, t.tw_inv, ps,
                           s.n, s.n2, s.n, s.r, ksteps, rtiles, 0u);
    }
    check_launch("ntt_columns_inverse_tiles");
}

void launch_crt(const Shape& s, const DeviceTables& t, const int8_t* rtiles, int64_t* words, const TileSet& ts,
                hipStream_t stream) {
    check_tiles(s, ts, "crt_gemm");
    const uint32_t blocks = (s.crt_rows / kTile) * ts.count;
    if (words_by_coefficient(s))
        hipLaunchKernelGGL(crt_gemm<true>, dim3(blocks), dim3(kGemmThreads), 0, stream, t.wcrt, rtiles, words,
                           s.n, s.n_pad, s.crt_rows, s.crt_k / kTK, s.crt_words, word_stride(s), ts);
    else
        hipLaunchKernelGGL(crt_gemm<false>, dim3(blocks), dim3(kGemmThreads), 0, stream, t.wcrt, rtiles, words,
                           s.n, s.n_pad, s.crt_rows, s.crt_k / kTK, s.crt_words, word_stride(s), ts);
    check_launch("crt_gemm");
}

void launch_reconstruct(const Shape& s, const DeviceTables& t, const int64_t* words, uint32_t* t_scratch,
                        uint32_t* c, const TileSet& ts, hipStream_t stream) {
    check_tiles(s, ts, "reconstruct");
    if (words_by_coefficient(s)) {
        // Barrett window: the four words of x from m up (m = floor(s/32) - 1, or 0 below s = 64)
        const uint32_t top_m = s.s_shift >= 64 ? s.s_shift / 32 - 1 : 0;
        WaveReconstructArgs args{s.n, word_stride(s), s.limbs, s.t_words, ts,
                                 top_m, s.s_shift - 32 * top_m, s.k_shift, s.mu};
        hipLaunchKernelGGL(reconstruct_waves, dim3(ts.count * 64), dim3(256), 0, stream, words, c, t.q, t.neg_p,
                           t.corr_t, t.corr_e, args);
        check_launch("reconstruct_waves");
        return;
    }
    if (s.t_words > 64) throw std::logic_error("reconstruct: the lane-per-coefficient kernel takes T of 64 words at most");
    ReconstructArgs args{s.n, s.n_pad, s.limbs, s.t_words, ts, s.s_shift, s.k_shift, s.mu};
    hipLaunchKernelGGL(reconstruct, dim3(ts.count * 4), dim3(64), 0, stream, words, t_scratch, c, t.q, t.neg_p,
                       t.corr_t, t.corr_e, args);
    check_launch("reconstruct");
}

GroupItem group_item(const Shape& s, const DeviceTables& t, const DeviceWork& w, const uint32_t* a,
                     const uint32_t* b, uint32_t* c) {
    GroupItem it{};
    it.n = s.n, it.n_pad = s.n_pad, it.limbs = s.limbs, it.r = s.r, it.r_pad = s.r_pad;
    it.conv_rows = s.conv_rows, it.conv_k = s.conv_k, it.crt_rows = s.crt_rows, it.crt_k = s.crt_k;
    it.crt_words = s.crt_words, it.ws = word_stride(s), it.t_words = s.t_words;
    it.n1 = s.n1, it.n2 = s.n2, it.rows = s.rows, it.row_width = s.row_width;
    it.s_shift = s.s_shift, it.k_shift = s.k_shift;
    it.top_m = s.s_shift >= 64 ? s.s_shift / 32 - 1 : 0;  // as launch_reconstruct
    it.top_e = s.s_shift - 32 * it.top_m;
    it.mu = s.mu;
    it.consts = t.consts, it.wconv = t.wconv, it.tw_fwd = t.tw_fwd, it.tw_inv = t.tw_inv, it.wcrt = t.wcrt;
    it.q = t.q, it.neg_p = t.neg_p, it.corr_t = t.corr_t, it.corr_e = t.corr_e;
    it.a = a, it.b = b, it.c = c;
    it.packed = w.packed, it.planes = w.planes, it.rtiles = w.rtiles, it.words = w.words, it.t_scratch = w.t_words;
    return it;
}

bool groupable(const Shape& s) {
    if (s.trivial) return false;
    if (s.rows >= 2 && (s.row_width % 256 != 0 || s.rows > 64)) return false;  // forward columns through LDS
    if (s.row_width % 16 != 0 || s.rows > 64) return false;                      // inverse columns through LDS
    return true;
}

namespace {

// Launches of one batch: pack, conversion, forward columns (one per row count),
// rows (one per block size n1, and the LDS kernel), inverse columns (one per
// row count), CRT GEMM (two word layouts), reconstruction (two kernels).
constexpr size_t kGroupLaunches = 2 + 6 + 5 + 7 + 2 + 2;

enum GroupKind { kPack, kConversion, kForward, kRows2048, kRowsProduct, kInverse, kCrt, kRecLane, kRecWave };

size_t group_item_bytes(size_t count) { return (count * sizeof(GroupItem) + 255) / 256 * 256; }
size_t group_table_words(size_t count) { return kGroupLaunches * (2 * count + 1); }

}  // namespace

size_t group_bytes(size_t count) { return group_item_bytes(count) + group_table_words(count) * sizeof(uint32_t); }

void launch_group(const std::vector<GroupItem>& items, void* dev_buffer, hipStream_t stream) {
    const uint32_t count = uint32_t(items.size());
    if (!count) return;
    const size_t item_bytes = group_item_bytes(count);
    const GroupItem* dev_items = static_cast<const GroupItem*>(dev_buffer);
    uint32_t* dev_tables = reinterpret_cast<uint32_t*>(static_cast<char*>(dev_buffer) + item_bytes);
    struct Launch {
        GroupKind kind;
        uint32_t variant;
        uint32_t offset, entries;
        uint64_t blocks;
        size_t lds;
    };
    std::vector<uint32_t> tables(group_table_words(count));
    std::vector<Launch> launches;
    size_t used = 0;
    // One launch over the items for which blocks(it) > 0.
    auto add = [&](GroupKind kind, uint32_t variant, auto blocks) {
        std::vector<uint32_t> sel;
        for (uint32_t i = 0; i < count; ++i)
            if (blocks(items[i]) > 0) sel.push_back(i);
        if (sel.empty()) return;
        if (launches.size() == kGroupLaunches) throw std::logic_error("launch_group: more launches than planned");
        uint32_t* start = tables.data() + used;
        uint32_t* item = start + sel.size() + 1;
        uint64_t total = 0;
        size_t lds = 0;
        for (size_t j = 0; j < sel.size(); ++j) {
            start[j] = uint32_t(total);
            item[j] = sel[j];
            total += blocks(items[sel[j]]);
            lds = std::max<size_t>(lds, size_t(2) * items[sel[j]].row_width * 4);
        }
        if (total >= (uint64_t(1) << 31)) throw std::invalid_argument("launch_group: batch too large for one grid");
        start[sel.size()] = uint32_t(total);
        launches.push_back(Launch{kind, variant, uint32_t(used), uint32_t(sel.size()), total, lds});
        used += 2 * sel.size() + 1;
    };
    auto ntiles = [](const GroupItem& it) { return uint64_t(it.n_pad / kTile); };
    auto rows_2048 = [](const GroupItem& it) { return it.row_width == kRowC && it.n1 <= 8; };
    add(kPack, 0, [&](const GroupItem& it) {  // as launch_pack: 2 * ntiles * 256 rows
        return 128 * ntiles(it) / pack_rows_per_wave(it.conv_k / kTK);
    });
    add(kConversion, 0, [&](const GroupItem& it) { return uint64_t(it.conv_rows / kTile) * 2 * ntiles(it); });
    for (uint32_t R = 2; R <= 64; R *= 2)
        add(kForward, R, [&](const GroupItem& it) {
            return it.rows == R ? uint64_t(it.row_width / 256) * it.r * 2 : 0;
        });
    for (uint32_t b = 1; b <= 8; b *= 2)
        add(kRows2048, b, [&](const GroupItem& it) {
            return rows_2048(it) && it.n1 == b ? uint64_t(it.rows) * it.r : 0;
        });
    add(kRowsProduct, 0, [&](const GroupItem& it) { return rows_2048(it) ? 0 : uint64_t(it.rows) * it.r; });
    for (uint32_t R = 1; R <= 64; R *= 2)
        add(kInverse, R, [&](const GroupItem& it) {
            return it.rows == R ? uint64_t(it.row_width / 16) * (it.r_pad / 16) : 0;
        });
    add(kCrt, 0, [&](const GroupItem& it) { return it.t_words > 64 ? 0 : uint64_t(it.crt_rows / kTile) * ntiles(it); });
    add(kCrt, 1, [&](const GroupItem& it) { return it.t_words > 64 ? uint64_t(it.crt_rows / kTile) * ntiles(it) : 0; });
    add(kRecLane, 0, [&](const GroupItem& it) { return it.t_words > 64 ? 0 : 4 * ntiles(it); });
    add(kRecWave, 0, [&](const GroupItem& it) { return it.t_words > 64 ? 64 * ntiles(it) : 0; });
    // the items and the work lists, one copy
    std::vector<char> upload(item_bytes + used * sizeof(uint32_t));
    std::memcpy(upload.data(), items.data(), count * sizeof(GroupItem));
    std::memcpy(upload.data() + item_bytes, tables.data(), used * sizeof(uint32_t));
    const hipError_t copied = hipMemcpy(dev_buffer, upload.data(), upload.size(), hipMemcpyHostToDevice);
    if (copied != hipSuccess) throw std::runtime_error(std::string("launch_group upload: ") + hipGetErrorString(copied));
    for (const Launch& L : launches) {
        const GroupLaunch g{dev_items, dev_tables + L.offset, dev_tables + L.offset + L.entries + 1, L.entries};
        const dim3 grid(uint32_t(L.blocks));
        switch (L.kind) {
            case kPack:
                hipLaunchKernelGGL(pack_group, grid, dim3(256), 0, stream, g);
                break;
            case kConversion:
                hipLaunchKernelGGL(conversion_group, grid, dim3(kGemmThreads), 0, stream, g);
                break;
            case kForward:
                switch (L.variant) {
#define PM_GFWD(R)                                                                      \
    case R:                                                                             \
        hipLaunchKernelGGL(forward_reg_group<R>, grid, dim3(256), 0, stream, g);         \
        break;
                    PM_GFWD(2) PM_GFWD(4) PM_GFWD(8) PM_GFWD(16) PM_GFWD(32) PM_GFWD(64)
#undef PM_GFWD
                }
                break;
            case kRows2048:
                switch (L.variant) {
#define PM_GROWS(B)                                                                     \
    case B:                                                                             \
        hipLaunchKernelGGL(rows_2048_group<B>, grid, dim3(256), 0, stream, g);           \
        break;
                    PM_GROWS(1) PM_GROWS(2) PM_GROWS(4) PM_GROWS(8)
#undef PM_GROWS
                }
                break;
            case kRowsProduct:
                hipLaunchKernelGGL(rows_product_group, grid, dim3(256), L.lds, stream, g);
                break;
            case kInverse:
                switch (L.variant) {
#define PM_GINV(R)                                                                      \
    case R:                                                                             \
        hipLaunchKernelGGL(inverse_tiles_group<R>, grid, dim3(256), 0, stream, g);       \
        break;
                    PM_GINV(1) PM_GINV(2) PM_GINV(4) PM_GINV(8) PM_GINV(16) PM_GINV(32) PM_GINV(64)
#undef PM_GINV
                }
                break;
            case kCrt:
                if (L.variant)
                    hipLaunchKernelGGL(crt_group<true>, grid, dim3(kGemmThreads), 0, stream, g);
                else
                    hipLaunchKernelGGL(crt_group<false>, grid, dim3(kGemmThreads), 0, stream, g);
                break;
            case kRecLane:
                hipLaunchKernelGGL(reconstruct_group, grid, dim3(64), 0, stream, g);
                break;
            case kRecWave:
                hipLaunchKernelGGL(reconstruct_waves_group, grid, dim3(256), 0, stream, g);
                break;
        }
        check_launch("grouped stage");
    }
}

void launch_zero(uint32_t* c, size_t words, hipStream_t stream) {
    if (!words) return;
    dim3 grid(uint32_t((words + 255) / 256));
    hipLaunchKernelGGL(fill_zero, grid, dim3(256), 0, stream, c, words);
    check_launch("fill_zero");
}

void launch_build_tables(const Shape& s, const TableBuild& t, hipStream_t stream) {
    const uint32_t r = s.r;
    launch_zero(reinterpret_cast<uint32_t*>(t.wconv), size_t(2) * s.conv_rows * s.conv_k / 4, stream);
    launch_zero(reinterpret_cast<uint32_t*>(t.wcrt), size_t(s.crt_rows) * s.crt_k / 4, stream);
    hipLaunchKernelGGL(table_beta, dim3(r), dim3(256), 0, stream, t.consts, t.rcp, t.beta, r, s.n2);
    check_launch("table_beta");
    const uint32_t cols = (s.n2 + 255) / 256;
    hipLaunchKernelGGL(table_twiddles, dim3(cols < 64 ? cols : 64, r), dim3(256), 0, stream, t.consts, t.rcp, t.psi,
                       t.tw_fwd, t.tw_inv, s.n2, s.log_n2);
    check_launch("table_twiddles");
    hipLaunchKernelGGL(table_conversion, dim3(r, 2), dim3(256), 0, stream, t.consts, t.rcp, t.beta, t.wconv,
                       s.conv_rows, s.conv_k, s.limbs);
    check_launch("table_conversion");
    hipLaunchKernelGGL(table_crt, dim3((r + 63) / 64), dim3(64), 0, stream, t.m_mod_q, t.rcp, t.wcrt, r, s.limbs,
                       s.crt_k);
    check_launch("table_crt");
}

// ---------------------------------------------------------------------------
// Static workspace of the allocation-free device entry point
// (pm_multiply_device_scratch).  The array is part of this code object: each
// GPU that loads the library holds its own copy, at its own address, until
// the process ends.  In host code `scratch` names only a host-side stand-in,
// so its device address is looked up once per device (init_scratch, called
// when the library makes a plan on that device) and only that address is
// handed to kernels.
// ---------------------------------------------------------------------------
namespace {
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kScratchDevices] = {};
hipError_t g_scratch_error[kScratchDevices] = {};
}  // namespace

void init_scratch() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScratchDevices) {
        (void)hipGetLastError();
        return;  // pm_multiply_device_scratch reports the device as unsupported
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev]) return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();  // keep the failure out of later launch checks
        g_scratch_error[dev] = e != hipSuccess ? e : hipErrorInvalidSymbol;
        return;
    }
    g_scratch_ptr[dev] = ptr;
}

void* scratch_address(int device, std::string* why) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (device < 0 || device >= kScratchDevices) {
        if (why) *why = "device " + std::to_string(device) + " is beyond the devices the static scratch supports";
        return nullptr;
    }
    if (!g_scratch_ptr[device] && why)
        *why = g_scratch_error[device] != hipSuccess
                   ? std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(g_scratch_error[device])
                   : "the static scratch was not set up on device " + std::to_string(device);
    return g_scratch_ptr[device];
}

size_t work_bytes(const Shape& s) { return carve_work(s, nullptr, nullptr); }

size_t carve_work(const Shape& s, uint8_t* base, DeviceWork* w) {
    const size_t packed = packed_operand_bytes(s);
    const size_t half = size_t(s.r_pad > scratch_rows(s) ? s.r_pad : scratch_rows(s)) * s.n_pad;
    const size_t planes = 2 * half * sizeof(uint32_t);
    const size_t words = word_sum_elements(s) * sizeof(int64_t);
    auto round = [](size_t b) { return (b + 255) / 256 * 256; };
    if (base && w) {
        w->packed = reinterpret_cast<int8_t*>(base);
        w->planes = reinterpret_cast<uint32_t*>(base + round(packed));
        w->rtiles = reinterpret_cast<int8_t*>(w->planes + half);
        w->t_words = w->planes;
        w->words = reinterpret_cast<int64_t*>(base + round(packed) + round(planes));
    }
    return round(packed) + round(planes) + round(words);
}

}  // namespace pm
