// Parameter-only plan construction (the work fherma_init does,
// candidates/ordered-streaming/solve.cu:900-1040), for the gfx950 pipeline:
//
//  primes      the reference's descending walk p = 1 (mod 2N), 2^30 <= p < 2^31,
//              until prod p > 4N(q-1)^2 (solve.cu:905-913).  When that walk runs
//              out of fields (wide q), the transform is split N = n1*n2 and the
//              walk is repeated with step 2*n2, doubling n1 until it succeeds.
//  conversion  byte weights w = lambda*256^j mod p as signed base-256 digits, so
//              signed-i8 MFMA can form sum_j byte_j(a) * 256^j mod p
//              (the reference's unsigned byte_weights, solve.cu:939-943).
//  NTT         merged-twist tables psi^bitrev(j) / psi^-bitrev(j) with Shoup
//              quotients (the reference twists separately, solve.cu:950-961).
//  CRT         M_k = P_k mod q for ARBITRARY q (the reference needs q = 2^W - gap,
//              solve.cu:916-925), the quotient reciprocals floor((2^64-1)/p_k)
//              (solve.cu:946), R = (-P) mod q, and the byte-convolution weight
//              matrix (solve.cu:964-979) in signed digits.
//  Barrett     mu = floor(2^(h+63)/q) for the final reduction of T mod q.
#include "plan.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <thread>

namespace pm {
namespace {

uint32_t pow_mod(uint32_t a, uint64_t e, uint32_t p) {
    uint64_t result = 1, base = a % p;
    for (; e; e >>= 1, base = base * base % p)
        if (e & 1) result = result * base % p;
    return uint32_t(result);
}

uint32_t inv_mod(uint32_t a, uint32_t p) { return pow_mod(a, p - 2, p); }

uint32_t mul_mod(uint32_t a, uint32_t b, uint32_t p) { return uint32_t(uint64_t(a) * b % p); }

uint32_t bit_reverse(uint32_t x, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

uint32_t shoup_quotient(uint32_t w, uint32_t p) { return uint32_t((uint64_t(w) << 32) / p); }

// Four signed base-256 digits e_d in [-128, 127] with sum e_d 256^d == c, for |c| < 2^30.
void signed_digits4(int64_t c, int8_t out[4]) {
    int64_t x = c;
    for (int d = 0; d < 4; ++d) {
        int64_t e = ((x + 128) & 255) - 128;
        out[d] = int8_t(e);
        x = (x - e) / 256;
    }
    if (x != 0) throw std::logic_error("plan: weight does not fit four signed digits");
}

// Signed base-256 digits of a non-negative multiword value: `count` digits in
// [-128, 127] (the last one absorbs the final carry).
std::vector<int8_t> signed_digits(const uint32_t* words, size_t nwords, size_t count) {
    std::vector<int8_t> out(count, 0);
    int carry = 0;
    for (size_t e = 0; e < count; ++e) {
        int byte = e / 4 < nwords ? int((words[e / 4] >> (8 * (e % 4))) & 255u) : 0;
        int x = byte + carry;
        if (x >= 128) {
            out[e] = int8_t(x - 256);
            carry = 1;
        } else {
            out[e] = int8_t(x);
            carry = 0;
        }
    }
    if (carry) throw std::logic_error("plan: signed digits overflow");
    return out;
}

thread_local unsigned t_plan_threads = 0;  // set_plan_threads

// Threads plan construction may use from this thread: the machine's, at most
// 16, and at most the caller's set_plan_threads share.
unsigned plan_threads() {
    unsigned threads = std::thread::hardware_concurrency();
    threads = threads ? std::min(threads, 16u) : 1u;
    return t_plan_threads ? std::min(threads, t_plan_threads) : threads;
}

// fn(begin, end) over [0, count) split across up to plan_threads() host
// threads (the calling thread takes the first share); serial below
// `min_share` per thread.  Every thread is joined before the first exception a
// share threw (std::bad_alloc, say) is rethrown; shares whose thread could not
// be started run on the calling thread.
template <class F>
void parallel_range(size_t count, size_t min_share, F&& fn) {
    unsigned threads = plan_threads();
    threads = unsigned(std::max<size_t>(1, std::min<size_t>(threads, count / std::max<size_t>(1, min_share))));
    std::vector<std::exception_ptr> errors(threads);
    auto share = [&fn, &errors, count, threads](unsigned t) {
        try {
            fn(count * t / threads, count * (t + 1) / threads);
        } catch (...) {
            errors[t] = std::current_exception();
        }
    };
    std::vector<std::thread> pool;
    pool.reserve(threads);
    unsigned started = 1;
    try {
        for (; started < threads; ++started) pool.emplace_back(share, started);
    } catch (...) {  // no more threads available
    }
    for (unsigned t = started; t < threads; ++t) share(t);
    share(0);
    for (auto& th : pool) th.join();
    for (const auto& e : errors)
        if (e) std::rethrow_exception(e);
}

// a mod p_k for each of `count` primes p_k < 2^31, with rcp_k =
// floor((2^64-1)/p_k): for x = rem*2^32 + word < p*2^32 the estimate
// hi64(x*rcp) is at most 2 short.  Words outer, primes inner, so the chains of
// different primes overlap instead of one prime's waiting on itself.
void mod_words_rcp(const uint32_t* a, size_t n, const uint32_t* p, const uint64_t* rcp, size_t count,
                   uint32_t* out) {
    std::vector<uint64_t> rem(count, 0);
    for (size_t i = n; i-- > 0;) {
        const uint64_t w = a[i];
        for (size_t k = 0; k < count; ++k) {
            const uint64_t x = (rem[k] << 32) | w;
            uint64_t r = x - uint64_t((unsigned __int128)x * rcp[k] >> 64) * p[k];
            r = std::min(r, r - p[k]);  // r - p wraps above r when r < p: no branch on random data
            r = std::min(r, r - p[k]);
            rem[k] = r;
        }
    }
    for (size_t k = 0; k < count; ++k) out[k] = uint32_t(rem[k]);
}

// M_g = (pq + q*t_g) / p_g for G exact divisions at once, without dividing:
// words of each quotient come from the low end as (word * p^-1 mod 2^32), each
// leaving a debt of hi32(quotient word * p) (plus a borrow) for the next word.
// The sums pq + q*t are formed in the same pass, and the G chains are
// independent, so they overlap.  pq and q have L words; each M gets L+1.
template <int G>
void exact_quotients(const uint32_t* pq, const uint32_t* q, size_t L, const uint32_t* t, const uint32_t* p,
                     const uint32_t* pinv, uint32_t* const* m) {
    uint64_t carry[G] = {}, debt[G] = {};
    for (size_t i = 0; i <= L; ++i) {
        const uint64_t qi = i < L ? q[i] : 0, pqi = i < L ? pq[i] : 0;
        for (int g = 0; g < G; ++g) {
            const uint64_t x = qi * t[g] + pqi + carry[g];  // < 2^63 + 2^33
            carry[g] = x >> 32;
            const int64_t s = int64_t(uint32_t(x)) - int64_t(debt[g]);
            const uint32_t w = uint32_t(s), word = w * pinv[g];
            m[g][i] = word;
            debt[g] = ((uint64_t(word) * p[g]) >> 32) + uint64_t(s < 0);
        }
    }
    for (int g = 0; g < G; ++g)
        if (carry[g] || debt[g]) throw std::logic_error("plan: CRT weight division not exact");
}

// x = x * w mod q in place, for x < q of L words and q of h >= 64 bits.  The
// quotient estimate floor(x_top * w / (q_top + 1)) from the top 64 bits of x
// and q never overshoots and falls short by at most 2; one pass forms
// x*w - e*q, and the rare shortfall costs a compare (decided at the top word)
// and a subtraction.
struct ModQ {
    const uint32_t* q;
    size_t L, h;
    uint64_t q_top;

    ModQ(const uint32_t* q_words, size_t limbs, size_t bits) : q(q_words), L(limbs), h(bits), q_top(top(q_words)) {}

    uint64_t top(const uint32_t* x) const {  // floor(x / 2^(h-64))
        const size_t sh = h - 64, w = sh / 32, b = sh % 32;
        unsigned __int128 v = 0;
        for (size_t i = 0; i < 3; ++i)
            if (w + i < L) v |= (unsigned __int128)x[w + i] << (32 * i);
        return uint64_t(v >> b);
    }

    bool at_least_q(const std::vector<uint32_t>& r) const {  // r has L+1 words
        if (r[L]) return true;
        for (size_t i = L; i-- > 0;)
            if (r[i] != q[i]) return r[i] > q[i];
        return true;
    }

    void mul_word(std::vector<uint32_t>& x, uint32_t w) const {  // x: L+1 words, x[L] == 0
        const uint64_t e = uint64_t((unsigned __int128)top(x.data()) * w / ((unsigned __int128)q_top + 1));
        uint64_t ca = 0, cb = 0, borrow = 0;
        for (size_t i = 0; i <= L; ++i) {
            const uint64_t a = (i < L ? uint64_t(x[i]) * w : 0) + ca;
            ca = a >> 32;
            const uint64_t b = (i < L ? uint64_t(q[i]) * e : 0) + cb;
            cb = b >> 32;
            const uint64_t d = uint64_t(uint32_t(a)) - uint32_t(b) - borrow;
            x[i] = uint32_t(d);
            borrow = d >> 63;
        }
        if (ca || cb || borrow) throw std::logic_error("plan: P mod q estimate overshot");
        for (int fix = 0; at_least_q(x); ++fix) {
            if (fix == 3) throw std::logic_error("plan: P mod q estimate too short");
            uint64_t br = 0;
            for (size_t i = 0; i <= L; ++i) {
                const uint64_t d = uint64_t(x[i]) - (i < L ? q[i] : 0) - br;
                x[i] = uint32_t(d);
                br = d >> 63;
            }
        }
    }
};

uint32_t newton_inverse(uint32_t p) {
    uint32_t inv = p;  // correct to 3 bits for odd p
    for (int i = 0; i < 4; ++i) inv *= 2u - p * inv;
    return inv;
}

}  // namespace

void set_plan_threads(unsigned threads) { t_plan_threads = threads; }

bool is_prime_u32(uint32_t p) {
    if (p < 2) return false;
    for (uint32_t a : {2u, 3u, 5u, 7u, 11u}) {
        if (p == a) return true;
        if (p % a == 0) return false;
    }
    uint32_t d = p - 1, shift = 0;
    while (!(d & 1)) {
        d >>= 1;
        ++shift;
    }
    // Bases {2,3,5,7,11} are deterministic below 2.15e12.
    for (uint32_t a : {2u, 3u, 5u, 7u, 11u}) {
        uint32_t x = pow_mod(a, d, p);
        if (x == 1 || x == p - 1) continue;
        bool witness = true;
        for (uint32_t j = 1; j < shift; ++j) {
            x = mul_mod(x, x, p);
            if (x == p - 1) {
                witness = false;
                break;
            }
        }
        if (witness) return false;
    }
    return true;
}

std::vector<int8_t> tile_256x64(const std::vector<int8_t>& src, uint32_t rows, uint32_t cols) {
    if (rows % 256 || cols % 64 || src.size() != size_t(rows) * cols)
        throw std::logic_error("plan: tiling shape mismatch");
    std::vector<int8_t> out(src.size());
    const uint32_t ksteps = cols / 64;
    for (uint32_t r = 0; r < rows; ++r)
        for (uint32_t ks = 0; ks < ksteps; ++ks) {
            const size_t dst = ((size_t(r / 256) * ksteps + ks) * 256 + (r % 256)) * 64;
            std::memcpy(out.data() + dst, src.data() + size_t(r) * cols + size_t(ks) * 64, 64);
        }
    return out;
}

std::vector<int8_t> untile_256x64(const std::vector<int8_t>& tiled, uint32_t rows, uint32_t cols) {
    if (rows % 256 || cols % 64 || tiled.size() != size_t(rows) * cols)
        throw std::logic_error("plan: tiling shape mismatch");
    std::vector<int8_t> out(tiled.size());
    const uint32_t ksteps = cols / 64;
    for (uint32_t r = 0; r < rows; ++r)
        for (uint32_t ks = 0; ks < ksteps; ++ks) {
            const size_t src = ((size_t(r / 256) * ksteps + ks) * 256 + (r % 256)) * 64;
            std::memcpy(out.data() + size_t(r) * cols + size_t(ks) * 64, tiled.data() + src, 64);
        }
    return out;
}

namespace {

// A positive number m * 2^e with m in [2^63, 2^64): a bound on a product that
// is kept rounded down (lower) or up (upper), so the walk can tell
// "product <= bound" and "product > bound" apart in O(1) per prime.
struct Approx {
    uint64_t m;
    int64_t e;
};

Approx approx_mul(Approx a, uint32_t w, bool up) {
    const unsigned __int128 x = (unsigned __int128)a.m * w;  // >= 2^63 for w >= 1
    int s = 0;
    while ((x >> s) >> 64) ++s;
    uint64_t m = uint64_t(x >> s);
    int64_t e = a.e + s;
    while (!(m >> 63)) {  // w small: shift up (exact)
        m <<= 1;
        --e;
    }
    if (up && s && (x & ((((unsigned __int128)1) << s) - 1))) {
        if (++m == 0) {
            m = uint64_t(1) << 63;
            ++e;
        }
    }
    return Approx{m, e};
}

int approx_compare(Approx a, Approx b) {
    if (a.e != b.e) return a.e < b.e ? -1 : 1;
    return a.m < b.m ? -1 : (a.m > b.m ? 1 : 0);
}

// The bound's value, rounded down (lower) and up (upper).
void approx_of(const Words& x, Approx* lower, Approx* upper) {
    const size_t bits = bit_length(x);
    const size_t sh = bits > 64 ? bits - 64 : 0;
    uint64_t m = 0;
    for (size_t i = 0; i < 64 && sh + i < bits; ++i)
        if ((x[(sh + i) / 32] >> ((sh + i) % 32)) & 1) m |= uint64_t(1) << i;
    bool rest = false;
    for (size_t i = 0; i < sh && !rest; ++i) rest = (x[i / 32] >> (i % 32)) & 1;
    int64_t e = int64_t(sh);
    while (!(m >> 63)) {  // fewer than 64 bits: exact
        m <<= 1;
        --e;
    }
    *lower = Approx{m, e};
    *upper = *lower;
    if (rest && ++upper->m == 0) *upper = Approx{uint64_t(1) << 63, e + 1};
}

}  // namespace

// Candidates are taken in the walk's descending order; their primality is
// tested a block at a time over host threads (blocks grow from 2048
// candidates).  The product is tracked by a lower and an upper bound (64-bit
// mantissas, rounded down and up): the walk continues while the upper bound is
// at most the bound's lower value and stops once the lower bound exceeds the
// bound's upper value.  Only if the two ever straddle it (the rounding error
// after r primes is below r * 2^-62 relative) is the exact product formed.
bool prime_walk(uint64_t step, const Words& bound, std::vector<uint32_t>* primes) {
    primes->clear();
    if (compare(Words{1}, bound) > 0) return true;
    Approx b_lo, b_hi;
    approx_of(bound, &b_lo, &b_hi);
    Approx lo{uint64_t(1) << 63, -63}, hi = lo;  // the product, 1
    const uint64_t floor = uint64_t(1) << 30;
    uint64_t candidate = ((uint64_t(1) << 31) - 2) / step * step + 1;
    std::vector<uint8_t> prime;
    Words product;  // exact product, only once the bounds straddle the bound
    size_t block = 2048;
    for (;;) {
        if (candidate < floor) return false;
        const size_t count = size_t(std::min<uint64_t>(block, (candidate - floor) / step + 1));
        prime.assign(count, 0);
        const uint64_t first = candidate;
        parallel_range(count, 256, [&](size_t i0, size_t i1) {
            for (size_t i = i0; i < i1; ++i) prime[i] = is_prime_u32(uint32_t(first - i * step));
        });
        for (size_t i = 0; i < count; ++i) {
            if (!prime[i]) continue;
            const uint32_t p = uint32_t(first - i * step);
            primes->push_back(p);
            if (product.empty()) {
                lo = approx_mul(lo, p, false);
                hi = approx_mul(hi, p, true);
                if (approx_compare(hi, b_lo) <= 0) continue;  // product <= bound
                if (approx_compare(lo, b_hi) > 0) return true;  // product > bound
                product = Words{1};
                for (uint32_t f : *primes) mul_word_inplace(product, f);
            } else {
                mul_word_inplace(product, p);
            }
            if (compare(product, bound) > 0) return true;
        }
        candidate = first - count * step;  // >= floor - step >= 0: step <= 2^30
        block = std::min<size_t>(block * 2, 32768);
    }
}

HostPlan build_plan(uint32_t n, uint32_t width, uint32_t limbs, const uint32_t* q_words, Tables tables,
                    uint32_t min_block) {
    if (!n || (n & (n - 1)) || n >= (1u << 29))
        throw std::invalid_argument("invalid point: N must be a power of two below 2^29");
    if (!width || !limbs || limbs != (width + 31) / 32)
        throw std::invalid_argument("invalid point: L must equal ceil(W/32)");
    if (!q_words) throw std::invalid_argument("invalid point: missing modulus");
    if (!min_block || (min_block & (min_block - 1)) || min_block > n || min_block > 2048)
        throw std::invalid_argument("invalid split: the minimum block must be a power of two <= min(N, 2048)");
    HostPlan plan;
    plan.tables = tables;
    const bool host_tables = tables == Tables::kHost;
    Shape& s = plan.shape;
    s.n = n;
    s.width = width;
    s.limbs = limbs;
    plan.q.assign(q_words, q_words + limbs);
    Words q(q_words, q_words + limbs);
    trim(q);
    if (is_zero(q)) throw std::invalid_argument("invalid point: q must be positive");
    s.qbits = uint32_t(bit_length(q));
    if (s.qbits == 1) {  // q == 1: the ring is zero
        s.trivial = true;
        return plan;
    }

    // ---- primes (solve.cu:905-913, generalized with the transform split) ----
    Words qm1 = sub(q, Words{1});
    Words bound = mul_word(mul(qm1, qm1), 4 * n);
    uint32_t n1 = min_block;
    bool found = false;
    for (; n1 <= n; n1 *= 2) {
        if (prime_walk(2 * uint64_t(n / n1), bound, &plan.primes)) {
            found = true;
            break;
        }
    }
    if (!found) throw InsufficientFields("no prime family of any transform split covers 4N(q-1)^2");
    s.n1 = n1;
    s.n2 = n / n1;
    s.log_n2 = 0;
    while ((1u << s.log_n2) < s.n2) ++s.log_n2;
    s.reference_walk = (n1 == 1);
    s.r = uint32_t(plan.primes.size());
    if (s.r >= 16384) throw InsufficientFields("the modulus needs more than 16383 primes");
    s.n_pad = (n + 255) / 256 * 256;
    s.r_pad = (s.r + 63) / 64 * 64;
    s.conv_k = (4 * limbs + 63) / 64 * 64;
    s.conv_rows = 4 * s.r_pad;
    s.crt_words = limbs + 4;
    s.crt_words_pad = (s.crt_words + 63) / 64 * 64;
    s.crt_rows = 4 * s.crt_words_pad;
    s.crt_k = 4 * s.r_pad;
    s.t_words = limbs + 2;
    s.row_width = n < 2048 ? n : 2048;
    if (s.row_width < n1) s.row_width = n1;
    s.rows = n / s.row_width;
    if (s.row_width > 2048) throw InsufficientFields("the transform split leaves blocks wider than 2048");
    if (s.rows > 8192) throw std::invalid_argument("invalid point: transform exceeds the supported size (N <= 2^24)");

    const uint32_t r = s.r, L = limbs, n2 = s.n2;
    plan.consts.assign(r, PrimeConst{0, 0, 0, 0});
    plan.psi.resize(size_t(r) * 2);
    if (host_tables) {
        plan.beta.resize(r);
        plan.wconv.assign(size_t(2) * s.conv_rows * s.conv_k, 0);
        plan.tw_fwd.assign(size_t(r) * n2 * 2, 0);
        plan.tw_inv.assign(size_t(r) * n2 * 2, 0);
    }

    // Per-prime constants (and, for host tables, each prime's rows): independent
    // per prime, spread over host threads.
    std::vector<std::string> prime_errors(plan_threads());
    std::atomic<unsigned> prime_part{0};
    parallel_range(r, host_tables ? 1 : 64, [&](size_t k0, size_t k1) {
        const unsigned part = prime_part.fetch_add(1);
        try {
            for (size_t k = k0; k < k1; ++k) {
                const uint32_t p = plan.primes[k];
                PrimeConst& c = plan.consts[k];
                c.p = p;
                c.pinv = newton_inverse(p);
                // psi: a primitive 2*n2-th root of unity (psi^n2 == -1), as solve.cu:951.
                uint32_t psi = 0;
                for (uint32_t base = 2; !psi; ++base) {
                    uint32_t root = pow_mod(base, (p - 1) / (2 * uint64_t(n2)), p);
                    if (pow_mod(root, n2, p) == p - 1) psi = root;
                }
                const uint32_t psi_inv = inv_mod(psi, p);
                plan.psi[2 * size_t(k)] = psi;
                plan.psi[2 * size_t(k) + 1] = psi_inv;
                if (!host_tables) continue;
                // P_k mod p_k = prod_{j != k} p_j mod p_k (for P_k^-1 mod p_k)
                uint64_t pk = 1;
                for (uint32_t j = 0; j < r; ++j)
                    if (j != k) pk = pk * (plan.primes[j] % p) % p;
                const uint32_t radix = uint32_t((uint64_t(1) << 32) % p);
                // b's scale: the base-case Montgomery product contributes R^-1 and the
                // unnormalized inverse transform n2, so scale b by n2^-1 * R * P_k^-1:
                // the inverse transform then yields v_k = x * P_k^-1 mod p_k directly.
                const uint32_t beta = mul_mod(mul_mod(inv_mod(n2 % p, p), radix, p), inv_mod(uint32_t(pk), p), p);
                plan.beta[k] = beta;
                // Weights carry an extra R so that one REDC of the byte sum lands on
                // lambda*x mod p with lambda = 1 (a) or beta (b).
                const uint32_t lambda[2] = {radix, mul_mod(radix, beta, p)};
                for (int op = 0; op < 2; ++op) {
                    uint32_t power = lambda[op];
                    int64_t sum = 0;
                    int8_t* base = plan.wconv.data() + (size_t(op) * s.conv_rows + 4 * size_t(k)) * s.conv_k;
                    for (uint32_t j = 0; j < 4 * L; ++j) {
                        int64_t centered = power > p / 2 ? int64_t(power) - p : int64_t(power);
                        int8_t dig[4];
                        signed_digits4(centered, dig);
                        for (int d = 0; d < 4; ++d) base[size_t(d) * s.conv_k + j] = dig[d];
                        sum += centered;
                        power = mul_mod(power, 256, p);
                    }
                    int64_t off = (128 * sum) % int64_t(p);
                    if (off < 0) off += p;
                    (op ? c.off_b : c.off_a) = uint32_t(off);
                }
                std::vector<uint32_t> pw(n2), pwi(n2);
                pw[0] = pwi[0] = 1;
                for (uint32_t e = 1; e < n2; ++e) {
                    pw[e] = mul_mod(pw[e - 1], psi, p);
                    pwi[e] = mul_mod(pwi[e - 1], psi_inv, p);
                }
                uint32_t* fwd = plan.tw_fwd.data() + size_t(k) * n2 * 2;
                uint32_t* inv = plan.tw_inv.data() + size_t(k) * n2 * 2;
                for (uint32_t j = 0; j < n2; ++j) {
                    uint32_t e = bit_reverse(j, s.log_n2);
                    fwd[2 * j] = pw[e];
                    fwd[2 * j + 1] = shoup_quotient(pw[e], p);
                    inv[2 * j] = pwi[e];
                    inv[2 * j + 1] = shoup_quotient(pwi[e], p);
                }
            }
        } catch (const std::exception& e) {
            prime_errors[part] = e.what();
        }
    });
    for (const auto& e : prime_errors)
        if (!e.empty()) throw std::logic_error(e);

    // ---- CRT constants for arbitrary q ----
    // Pq = P mod q by one multiply-by-word-mod-q per prime.
    Words pq{1};
    if (s.qbits >= 64) {
        std::vector<uint32_t> qw(limbs, 0), x(limbs + 1, 0);
        std::copy(q.begin(), q.end(), qw.begin());
        const ModQ mq(qw.data(), limbs, s.qbits);
        x[0] = 1;
        for (uint32_t k = 0; k < r; ++k) mq.mul_word(x, plan.primes[k]);
        pq.assign(x.begin(), x.begin() + limbs);
        trim(pq);
    } else {
        for (uint32_t k = 0; k < r; ++k) pq = mul_word_mod(pq, plan.primes[k], q);
    }
    Words neg_p = is_zero(pq) ? Words{0} : sub(q, pq);
    plan.neg_p_mod_q.assign(L, 0);
    for (size_t i = 0; i < neg_p.size() && i < L; ++i) plan.neg_p_mod_q[i] = neg_p[i];

    plan.m_mod_q.assign(size_t(r) * L, 0);
    plan.reciprocal.resize(r);
    for (uint32_t k = 0; k < r; ++k) plan.reciprocal[k] = UINT64_MAX / plan.primes[k];
    // M_k for every prime: independent O(L) work per prime, spread over host
    // threads (each writes its own rows of m_mod_q and sums them into its own
    // 64-bit per-word accumulators: r < 2^14 terms below 2^32 each).
    std::vector<uint32_t> qw(L, 0), pqw(L, 0);
    for (size_t i = 0; i < q.size() && i < L; ++i) qw[i] = q[i];
    for (size_t i = 0; i < pq.size() && i < L; ++i) pqw[i] = pq[i];
    const unsigned threads = plan_threads();
    std::vector<std::vector<uint64_t>> partial(threads, std::vector<uint64_t>(L + 1, 0));
    std::vector<std::string> errors(threads);
    std::atomic<unsigned> next_part{0};
    parallel_range(r, std::max<size_t>(1, (size_t(1) << 16) / (L + 1)), [&](size_t k0, size_t k1) {
        const unsigned part = next_part.fetch_add(1);
        std::vector<uint64_t>& acc = partial[part];
        constexpr int kGroup = 4;
        std::vector<uint32_t> q_res(k1 - k0), pq_res(k1 - k0), mbuf(kGroup * (L + 1));
        // M < q (1 + 1/p_k): at most one subtraction of q; then store and sum
        auto finish = [&](size_t k, uint32_t* m) {
            bool ge = true;  // M >= q; equal counts
            if (!m[L]) {
                for (size_t i = L; i-- > 0;) {
                    if (m[i] != qw[i]) {
                        ge = m[i] > qw[i];
                        break;
                    }
                }
            }
            if (ge) {
                uint64_t borrow = 0;
                for (size_t i = 0; i <= L; ++i) {
                    const uint64_t d = uint64_t(m[i]) - (i < L ? qw[i] : 0) - borrow;
                    m[i] = uint32_t(d);
                    borrow = d >> 63;
                }
            }
            if (m[L]) throw std::logic_error("plan: CRT weight not reduced");
            for (size_t i = 0; i < L; ++i) {
                plan.m_mod_q[k * L + i] = m[i];
                acc[i] += m[i];
            }
        };
        try {
            mod_words_rcp(qw.data(), L, plan.primes.data() + k0, plan.reciprocal.data() + k0, k1 - k0, q_res.data());
            mod_words_rcp(pqw.data(), L, plan.primes.data() + k0, plan.reciprocal.data() + k0, k1 - k0,
                          pq_res.data());
            // M_k = (Pq + q*t)/p_k with t = -Pq * q^-1 mod p_k: exact, and
            // congruent to Pq * p_k^-1 = P_k (mod q).  O(L) per prime, kGroup
            // primes per pass.
            size_t ks[kGroup];
            uint32_t ts[kGroup], ps[kGroup], pinvs[kGroup];
            uint32_t* ms[kGroup];
            for (int g = 0; g < kGroup; ++g) ms[g] = mbuf.data() + g * (L + 1);
            int filled = 0;
            auto flush = [&]() {
                if (filled == kGroup) {
                    exact_quotients<kGroup>(pqw.data(), qw.data(), L, ts, ps, pinvs, ms);
                } else {
                    for (int g = 0; g < filled; ++g)
                        exact_quotients<1>(pqw.data(), qw.data(), L, ts + g, ps + g, pinvs + g, ms + g);
                }
                for (int g = 0; g < filled; ++g) finish(ks[g], ms[g]);
                filled = 0;
            };
            for (size_t k = k0; k < k1; ++k) {
                const uint32_t p = plan.primes[k];
                const uint32_t q_mod = q_res[k - k0];
                if (q_mod == 0) {
                    // p_k divides q: prod_{j != k} p_j mod q directly.
                    Words mk{1};
                    for (uint32_t j = 0; j < r; ++j)
                        if (j != k) mk = mul_word_mod(mk, plan.primes[j], q);
                    std::vector<uint32_t> m(L + 1, 0);
                    for (size_t i = 0; i < mk.size() && i < L; ++i) m[i] = mk[i];
                    finish(k, m.data());
                    continue;
                }
                ks[filled] = k;
                ts[filled] = (p - mul_mod(pq_res[k - k0], inv_mod(q_mod, p), p)) % p;
                ps[filled] = p;
                pinvs[filled] = plan.consts[k].pinv;
                if (++filled == kGroup) flush();
            }
            flush();
        } catch (const std::exception& e) {
            errors[part] = e.what();
        }
    });
    for (const auto& e : errors)
        if (!e.empty()) throw std::logic_error(e);
    Words sum_m(L + 2, 0);
    {
        unsigned __int128 carry = 0;
        for (size_t i = 0; i <= L + 1; ++i) {
            unsigned __int128 x = carry;
            if (i <= L)
                for (const auto& acc : partial) x += acc[i];
            sum_m[i] = uint32_t(x);
            carry = x >> 32;
        }
        if (carry) throw std::logic_error("plan: CRT weight sum overflow");
        trim(sum_m);
    }
    unsigned __int128 sum_rcp = 0;
    for (uint32_t k = 0; k < r; ++k) sum_rcp += plan.reciprocal[k];

    // Signed-byte offset corrections: the GEMM sees residue bytes minus 128.
    Words corr = mul_word(mul_word(sum_m, 0x01010101u), 128);
    plan.corr_t.assign(s.t_words, 0);
    for (size_t i = 0; i < corr.size(); ++i) {
        if (i >= s.t_words) {
            if (corr[i]) throw std::logic_error("plan: correction exceeds T width");
            continue;
        }
        plan.corr_t[i] = corr[i];
    }
    Words rcp_words{uint32_t(sum_rcp), uint32_t(sum_rcp >> 32), uint32_t(sum_rcp >> 64), uint32_t(sum_rcp >> 96)};
    trim(rcp_words);
    Words corr_e = mul_word(mul_word(rcp_words, 0x01010101u), 128);
    plan.corr_e.assign(4, 0);
    for (size_t i = 0; i < corr_e.size(); ++i) {
        if (i >= 4) throw std::logic_error("plan: quotient correction too wide");
        plan.corr_e[i] = corr_e[i];
    }

    // CRT weights: row c = output byte column, col 4k+d = byte d of v_k.
    // T rows: c in [0, 4(L+1)), weight = digit_{c-d}(M_k) (4L+1 signed digits).
    // Quotient rows: c = 4(L+1) + e, weight = digit_{e-d}(rcp_k) (6 signed digits).
    const size_t md = 4 * size_t(L) + 1;
    if (host_tables) plan.wcrt.assign(size_t(s.crt_rows) * s.crt_k, 0);
    for (uint32_t k = 0; k < r && host_tables; ++k) {
        std::vector<int8_t> dm = signed_digits(plan.m_mod_q.data() + size_t(k) * L, L, md);
        uint32_t rw[2] = {uint32_t(plan.reciprocal[k]), uint32_t(plan.reciprocal[k] >> 32)};
        std::vector<int8_t> dr = signed_digits(rw, 2, 6);
        for (uint32_t d = 0; d < 4; ++d) {
            const size_t col = 4 * size_t(k) + d;
            for (size_t e = 0; e < md; ++e) {
                if (!dm[e]) continue;
                plan.wcrt[(e + d) * s.crt_k + col] = dm[e];
            }
            for (size_t e = 0; e < 6; ++e) {
                if (!dr[e]) continue;
                plan.wcrt[(4 * (size_t(L) + 1) + e + d) * s.crt_k + col] = dr[e];
            }
        }
    }

    // Barrett constants: T < 2^45 q; Ts = floor(T/2^s) with s = max(h-16, 0),
    // Q = floor(Ts*mu / 2^(h-s+63)) with mu = floor(2^(h+63)/q) gives
    // Q in {floor(T/q) - 1, floor(T/q)}.
    const uint32_t h = s.qbits - 1;
    s.s_shift = h >= 16 ? h - 16 : 0;
    s.k_shift = h - s.s_shift + 63;
    s.mu = div_u64_quotient(power_of_two(size_t(h) + 63), q);
    return plan;
}

}  // namespace pm
