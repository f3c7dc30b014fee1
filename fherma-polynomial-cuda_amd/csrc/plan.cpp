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
#include <cstring>
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

uint32_t newton_inverse(uint32_t p) {
    uint32_t inv = p;  // correct to 3 bits for odd p
    for (int i = 0; i < 4; ++i) inv *= 2u - p * inv;
    return inv;
}

}  // namespace

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

bool prime_walk(uint64_t step, const Words& bound, std::vector<uint32_t>* primes) {
    primes->clear();
    Words product{1};
    uint64_t candidate = ((uint64_t(1) << 31) - 2) / step * step + 1;
    while (compare(product, bound) <= 0) {
        if (candidate < (uint64_t(1) << 30)) return false;
        if (is_prime_u32(uint32_t(candidate))) {
            primes->push_back(uint32_t(candidate));
            product = mul_word(product, uint32_t(candidate));
        }
        candidate -= step;
    }
    return true;
}

HostPlan build_plan(uint32_t n, uint32_t width, uint32_t limbs, const uint32_t* q_words, Tables tables) {
    if (!n || (n & (n - 1)) || n >= (1u << 29))
        throw std::invalid_argument("invalid point: N must be a power of two below 2^29");
    if (!width || !limbs || limbs != (width + 31) / 32)
        throw std::invalid_argument("invalid point: L must equal ceil(W/32)");
    if (!q_words) throw std::invalid_argument("invalid point: missing modulus");
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
    uint32_t n1 = 1;
    bool found = false;
    for (; n1 <= n; n1 *= 2) {
        if (prime_walk(2 * uint64_t(n / n1), bound, &plan.primes)) {
            found = true;
            break;
        }
    }
    if (!found) throw std::runtime_error("insufficient independent NTT fields");
    s.n1 = n1;
    s.n2 = n / n1;
    s.log_n2 = 0;
    while ((1u << s.log_n2) < s.n2) ++s.log_n2;
    s.reference_walk = (n1 == 1);
    s.r = uint32_t(plan.primes.size());
    if (s.r >= 16384) throw std::runtime_error("modulus too wide: more than 16383 NTT fields needed");
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
    if (s.row_width > 2048) throw std::runtime_error("insufficient independent NTT fields");
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

    for (uint32_t k = 0; k < r; ++k) {
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

    // ---- CRT constants for arbitrary q ----
    // Pq = P mod q by one multiply-by-word-mod-q per prime.
    Words pq{1};
    for (uint32_t k = 0; k < r; ++k) pq = mul_word_mod(pq, plan.primes[k], q);
    Words neg_p = is_zero(pq) ? Words{0} : sub(q, pq);
    plan.neg_p_mod_q.assign(L, 0);
    for (size_t i = 0; i < neg_p.size() && i < L; ++i) plan.neg_p_mod_q[i] = neg_p[i];

    plan.m_mod_q.assign(size_t(r) * L, 0);
    plan.reciprocal.resize(r);
    // M_k for every prime: independent O(L) work per prime, spread over a few
    // host threads (each writes its own rows of m_mod_q).
    auto m_range = [&](uint32_t k0, uint32_t k1, std::string* error) {
        try {
            for (uint32_t k = k0; k < k1; ++k) {
                const uint32_t p = plan.primes[k];
                Words m;
                const uint32_t q_mod = mod_word(q, p);
                if (q_mod != 0) {
                    // M_k = (Pq + q*t)/p_k with t = -Pq * q^-1 mod p_k: exact, and
                    // congruent to Pq * p_k^-1 = P_k (mod q).  O(L) per prime.
                    const uint32_t pq_mod = mod_word(pq, p);
                    const uint32_t t = (p - mul_mod(pq_mod, inv_mod(q_mod, p), p)) % p;
                    uint32_t rem = 0;
                    m = divmod_word(add(pq, mul_word(q, t)), p, &rem);
                    if (rem) throw std::logic_error("plan: CRT weight division not exact");
                    if (compare(m, q) >= 0) m = sub(m, q);
                } else {
                    // p_k divides q: prod_{j != k} p_j mod q directly.
                    m = Words{1};
                    for (uint32_t j = 0; j < r; ++j)
                        if (j != k) m = mul_word_mod(m, plan.primes[j], q);
                }
                for (size_t i = 0; i < m.size() && i < L; ++i) plan.m_mod_q[size_t(k) * L + i] = m[i];
            }
        } catch (const std::exception& e) {
            *error = e.what();
        }
    };
    {
        const uint64_t work = uint64_t(r) * L;
        unsigned threads = std::thread::hardware_concurrency();
        threads = threads ? std::min(threads, 16u) : 1u;
        if (work < (uint64_t(1) << 18)) threads = 1;
        threads = std::min<unsigned>(threads, r);
        std::vector<std::string> errors(threads);
        std::vector<std::thread> pool;
        for (unsigned t = 1; t < threads; ++t)
            pool.emplace_back(m_range, uint32_t(uint64_t(r) * t / threads), uint32_t(uint64_t(r) * (t + 1) / threads),
                              &errors[t]);
        m_range(0, uint32_t(uint64_t(r) / threads), &errors[0]);
        for (auto& th : pool) th.join();
        for (const auto& e : errors)
            if (!e.empty()) throw std::logic_error(e);
    }
    Words sum_m{0};
    unsigned __int128 sum_rcp = 0;
    for (uint32_t k = 0; k < r; ++k) {
        Words m(plan.m_mod_q.begin() + size_t(k) * L, plan.m_mod_q.begin() + size_t(k + 1) * L);
        trim(m);
        sum_m = add(sum_m, m);
        plan.reciprocal[k] = UINT64_MAX / plan.primes[k];
        sum_rcp += plan.reciprocal[k];
    }

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
