// Minimal unsigned host big integers (little-endian 32-bit words) for the
// parameter-only plan: prime products, P_k mod q, reciprocals.  Only plan
// construction uses this; nothing here runs per multiplication.
//
// Plays the role of the reference's host helper candidates/ordered-streaming/big_unsigned.h
// (trim/compare/subtract/times_word/multiply/divide_word/modulo_word/modulo),
// written independently; tests check it against that header compiled as-is.
#pragma once

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace pm {

using Words = std::vector<uint32_t>;

inline void trim(Words& a) {
    while (a.size() > 1 && a.back() == 0) a.pop_back();
    if (a.empty()) a.push_back(0);
}

inline bool is_zero(const Words& a) {
    for (uint32_t w : a)
        if (w) return false;
    return true;
}

inline uint64_t low64(const Words& a) {
    uint64_t v = a.size() > 0 ? a[0] : 0;
    if (a.size() > 1) v |= uint64_t(a[1]) << 32;
    return v;
}

inline size_t bit_length(const Words& a) {
    for (size_t i = a.size(); i-- > 0;)
        if (a[i]) return i * 32 + (32 - __builtin_clz(a[i]));
    return 0;
}

inline int compare(const Words& a, const Words& b) {
    size_t n = std::max(a.size(), b.size());
    for (size_t i = n; i-- > 0;) {
        uint32_t x = i < a.size() ? a[i] : 0, y = i < b.size() ? b[i] : 0;
        if (x != y) return x < y ? -1 : 1;
    }
    return 0;
}

inline Words add(const Words& a, const Words& b) {
    size_t n = std::max(a.size(), b.size());
    Words out(n + 1, 0);
    uint64_t carry = 0;
    for (size_t i = 0; i < n; ++i) {
        uint64_t t = carry + (i < a.size() ? a[i] : 0) + (i < b.size() ? b[i] : 0);
        out[i] = uint32_t(t);
        carry = t >> 32;
    }
    out[n] = uint32_t(carry);
    trim(out);
    return out;
}

// a - b, requires a >= b.
inline Words sub(const Words& a, const Words& b) {
    if (compare(a, b) < 0) throw std::logic_error("bigint: negative difference");
    Words out(a.size(), 0);
    int64_t borrow = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        int64_t t = int64_t(a[i]) - (i < b.size() ? int64_t(b[i]) : 0) - borrow;
        borrow = t < 0;
        out[i] = uint32_t(t + (borrow ? (int64_t(1) << 32) : 0));
    }
    trim(out);
    return out;
}

inline Words mul_word(const Words& a, uint32_t w) {
    Words out(a.size() + 1, 0);
    uint64_t carry = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        uint64_t t = uint64_t(a[i]) * w + carry;
        out[i] = uint32_t(t);
        carry = t >> 32;
    }
    out[a.size()] = uint32_t(carry);
    trim(out);
    return out;
}

// a *= w in place (w != 0 keeps a trimmed).
inline void mul_word_inplace(Words& a, uint32_t w) {
    uint64_t carry = 0;
    for (auto& x : a) {
        const uint64_t t = uint64_t(x) * w + carry;
        x = uint32_t(t);
        carry = t >> 32;
    }
    if (carry) a.push_back(uint32_t(carry));
}

inline Words mul(const Words& a, const Words& b) {
    Words out(a.size() + b.size(), 0);
    for (size_t i = 0; i < a.size(); ++i) {
        uint64_t carry = 0;
        for (size_t j = 0; j < b.size(); ++j) {
            uint64_t t = uint64_t(a[i]) * b[j] + out[i + j] + carry;
            out[i + j] = uint32_t(t);
            carry = t >> 32;
        }
        out[i + b.size()] = uint32_t(carry);
    }
    trim(out);
    return out;
}

inline Words shl(const Words& a, size_t bits) {
    size_t words = bits / 32, shift = bits % 32;
    Words out(a.size() + words + 1, 0);
    for (size_t i = 0; i < a.size(); ++i) {
        uint64_t v = uint64_t(a[i]) << shift;
        out[i + words] |= uint32_t(v);
        out[i + words + 1] |= uint32_t(v >> 32);
    }
    trim(out);
    return out;
}

inline Words power_of_two(size_t bits) {
    Words out(bits / 32 + 1, 0);
    out[bits / 32] = 1u << (bits % 32);
    return out;
}

// Quotient and remainder by one word.
inline Words divmod_word(const Words& a, uint32_t d, uint32_t* remainder) {
    if (!d) throw std::logic_error("bigint: division by zero");
    Words out(a.size(), 0);
    uint64_t rem = 0;
    for (size_t i = a.size(); i-- > 0;) {
        uint64_t t = (rem << 32) | a[i];
        out[i] = uint32_t(t / d);
        rem = t % d;
    }
    if (remainder) *remainder = uint32_t(rem);
    trim(out);
    return out;
}

inline uint32_t mod_word(const Words& a, uint32_t d) {
    uint64_t rem = 0;
    for (size_t i = a.size(); i-- > 0;) rem = ((rem << 32) | a[i]) % d;
    return uint32_t(rem);
}

// Bit-serial long division; only used on rare or tiny paths (quotients of a few
// dozen bits, or single fall-back residues), never per multiplication.
inline Words divmod(const Words& a, const Words& d, Words* remainder) {
    if (is_zero(d)) throw std::logic_error("bigint: division by zero");
    size_t na = bit_length(a);
    Words q(a.size() + 1, 0), r{0};
    for (size_t bit = na; bit-- > 0;) {
        r = shl(r, 1);
        if ((a[bit / 32] >> (bit % 32)) & 1) r[0] |= 1;
        if (compare(r, d) >= 0) {
            r = sub(r, d);
            q[bit / 32] |= 1u << (bit % 32);
        }
    }
    trim(q);
    trim(r);
    if (remainder) *remainder = r;
    return q;
}

// t - w*q for t >= w*q (one word-times-bigint subtraction, O(len)).
inline Words sub_mul_word(const Words& t, const Words& q, uint32_t w) {
    return sub(t, mul_word(q, w));
}

// floor(t / q) and t mod q for t < q * 2^32: the quotient is one word.  The
// estimate divides t's top 64 bits (aligned to q's top 32 bits) by q's top
// word plus one, which never overshoots and falls short by at most a few
// units; exact subtractions finish it.  O(len(q)) per call.
inline uint32_t divmod_small(const Words& t, const Words& q, Words* remainder) {
    size_t h = bit_length(q);
    if (h == 0) throw std::logic_error("bigint: division by zero");
    if (h <= 32) {
        uint64_t tv = low64(t);
        if (bit_length(t) > 64) throw std::logic_error("bigint: quotient exceeds one word");
        uint64_t qv = q[0];
        if (remainder) *remainder = Words{uint32_t(tv % qv)};
        uint64_t quotient = tv / qv;
        if (quotient >> 32) throw std::logic_error("bigint: quotient exceeds one word");
        return uint32_t(quotient);
    }
    size_t shift = h - 32;
    auto bits_at = [](const Words& a, size_t from, size_t count) {
        uint64_t v = 0;
        for (size_t i = 0; i < count; ++i) {
            size_t bit = from + i;
            if (bit / 32 < a.size() && ((a[bit / 32] >> (bit % 32)) & 1)) v |= uint64_t(1) << i;
        }
        return v;
    };
    uint64_t qt = bits_at(q, shift, 32);
    if (bit_length(t) > h + 32) throw std::logic_error("bigint: quotient exceeds one word");
    uint64_t tt = bits_at(t, shift, 64);
    uint64_t estimate = tt / (qt + 1);
    if (estimate >> 32) estimate = 0xffffffffu;
    Words r = sub_mul_word(t, q, uint32_t(estimate));
    while (compare(r, q) >= 0) {
        r = sub(r, q);
        ++estimate;
    }
    if (estimate >> 32) throw std::logic_error("bigint: quotient exceeds one word");
    if (remainder) *remainder = r;
    return uint32_t(estimate);
}

// x * w mod q for x < q.
inline Words mul_word_mod(const Words& x, uint32_t w, const Words& q) {
    Words r;
    divmod_small(mul_word(x, w), q, &r);
    return r;
}

// floor(a / q) for a < q * 2^64 (a two-word quotient), by two one-word steps.
inline uint64_t div_u64_quotient(const Words& a, const Words& q) {
    // high part: floor(a / 2^32) < q * 2^32
    Words hi(a.begin() + std::min<size_t>(1, a.size()), a.end());
    if (hi.empty()) hi.push_back(0);
    Words r1;
    uint32_t q1 = divmod_small(hi, q, &r1);
    Words a2 = shl(r1, 32);
    a2[0] |= a.empty() ? 0 : a[0];
    uint32_t q0 = divmod_small(a2, q, nullptr);
    return (uint64_t(q1) << 32) | q0;
}

}  // namespace pm
