// Parameter-only plan for c = a*b mod (X^N + 1, q): everything fherma_init
// computes before any operand is seen (candidates/ordered-streaming/solve.cu:900-1040),
// re-designed for the gfx950 kernels in kernels.hip.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "bigint.hpp"

namespace pm {

// No set of NTT primes the plan can use covers 4N(q-1)^2 (the reference's
// "insufficient independent NTT fields", solve.cu:910); PM_ERR_INSUFFICIENT_FIELDS.
struct InsufficientFields : std::runtime_error {
    explicit InsufficientFields(const std::string& what)
        : std::runtime_error("insufficient independent NTT fields: " + what) {}
};

// Per-prime constants read by the kernels.
struct PrimeConst {
    uint32_t p;      // NTT prime, 2^30 <= p < 2^31
    uint32_t pinv;   // p^-1 mod 2^32 (subtractive Montgomery)
    uint32_t off_a;  // conversion offset for operand a (signed-byte correction), mod p
    uint32_t off_b;  // same for operand b
};

struct Shape {
    uint32_t n = 0, width = 0, limbs = 0;
    // Transform split: N = n1 * n2.  The NTT of length n2 runs over the n1
    // interleaved sub-sequences a_{t + n1*u}; the base case multiplies degree<n1
    // polynomials mod (X^n1 - zeta).  n1 = 1 is the plain negacyclic NTT.
    uint32_t n1 = 1, n2 = 1, log_n2 = 0;
    uint32_t r = 0;          // primes
    uint32_t n_pad = 0;      // N padded to the 256-coefficient GEMM tile: row stride of planes/words
    uint32_t r_pad = 0;      // primes padded to a multiple of 64 (256 GEMM rows)
    uint32_t conv_k = 0;     // conversion GEMM K: 4L bytes padded to 64
    uint32_t conv_rows = 0;  // conversion GEMM rows: 4 digits per (padded) prime
    uint32_t crt_words = 0;  // CRT GEMM output words: L+1 sum words + 3 quotient words
    uint32_t crt_words_pad = 0;
    uint32_t crt_rows = 0;   // 4 bytes per (padded) word
    uint32_t crt_k = 0;      // 4 bytes per (padded) prime
    uint32_t t_words = 0;    // L+2: width of T = sum v_k M_k + K R
    uint32_t row_width = 1;  // NTT row length C (flat elements); rows = N / C
    uint32_t rows = 1;
    uint32_t qbits = 0;      // bit length of q
    uint32_t s_shift = 0;    // Barrett: Ts = floor(T / 2^s_shift)
    uint32_t k_shift = 0;    // Barrett: Q = floor(Ts * mu / 2^k_shift)
    uint64_t mu = 0;         // floor(2^(qbits-1+63) / q)
    bool trivial = false;    // q == 1: every output is zero
    bool reference_walk = true;  // primes are exactly fherma_init's walk (n1 == 1)
};

// Which tables build_plan fills on the host.  kDevice leaves the prime-set
// tables (wconv, tw_fwd, tw_inv, beta, conversion offsets) and the CRT weight
// matrix wcrt empty: the device builds them from primes, psi, reciprocal and
// m_mod_q (kernels.hpp, launch_build_tables).
enum class Tables { kHost, kDevice };

struct HostPlan {
    Shape shape;
    Tables tables = Tables::kHost;
    std::vector<uint32_t> primes;
    std::vector<PrimeConst> consts;   // off_a/off_b are 0 under Tables::kDevice
    std::vector<uint32_t> psi;        // [r][2]  psi (primitive 2*n2-th root), psi^-1
    std::vector<int8_t> wconv;        // [2][conv_rows][conv_k]
    std::vector<uint32_t> tw_fwd;     // [r][n2][2]  (psi^bitrev(j), Shoup quotient)
    std::vector<uint32_t> tw_inv;     // [r][n2][2]  (psi^-bitrev(j), Shoup quotient)
    std::vector<int8_t> wcrt;         // [crt_rows][crt_k]
    std::vector<uint32_t> q;          // [L]
    std::vector<uint32_t> neg_p_mod_q;  // [L]  R = (-P) mod q
    std::vector<uint32_t> corr_t;     // [t_words]
    std::vector<uint32_t> corr_e;     // [4]
    // kept for inspection by tests
    std::vector<uint32_t> m_mod_q;    // [r][L]  M_k = P_k mod q
    std::vector<uint64_t> reciprocal; // [r]     floor((2^64-1)/p_k)
    std::vector<uint32_t> beta;       // [r]     operand-b scale folded into its weights
};

// Validation of the point, as fherma_init does (solve.cu:903-904): N a power of
// two below 2^29, L == ceil(W/32), q of exactly L limbs' storage, 1 <= q.
// Throws std::invalid_argument / std::runtime_error with a message.
// The transform split n1 is the smallest power of two >= min_block whose prime
// family covers 4N(q-1)^2 (min_block = 1 is the plain choice; larger values pin
// each base-case kernel at small sizes).
HostPlan build_plan(uint32_t n, uint32_t width, uint32_t limbs, const uint32_t* q_words,
                    Tables tables = Tables::kHost, uint32_t min_block = 1);

// fherma_init's prime walk, generalized: descending p = 1 (mod step) from
// floor((2^31-2)/step)*step + 1, primes only, while prod(p) <= bound, failing
// below 2^30.  step = 2N reproduces solve.cu:905-913 exactly.
bool prime_walk(uint64_t step, const Words& bound, std::vector<uint32_t>* primes);
// Host threads that plan construction started from the calling thread may use
// (0: as many as the machine has, at most 16).  A server's workers each take a
// share, so that concurrent plan builds do not oversubscribe the host.
void set_plan_threads(unsigned threads);

bool is_prime_u32(uint32_t p);

// GEMM operand tiling: a row-major [rows][cols] byte matrix (rows a multiple
// of 256, cols a multiple of 64) as [rows/256][cols/64][256][64], so that every
// (256-row tile, 64-byte K step) block is one contiguous 16 KB chunk.
std::vector<int8_t> tile_256x64(const std::vector<int8_t>& src, uint32_t rows, uint32_t cols);
// Inverse of tile_256x64.
std::vector<int8_t> untile_256x64(const std::vector<int8_t>& tiled, uint32_t rows, uint32_t cols);

}  // namespace pm
