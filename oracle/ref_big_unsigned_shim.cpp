// TEST INFRASTRUCTURE ONLY.  Builds the reference's own host big-integer
// helpers (candidates/ordered-streaming/big_unsigned.h, compiled where it lies
// under the reference tree; see oracle/build_oracle.py) into oracle/_ref/ and
// exposes them through a C ABI so that tests can check the oracle and the host
// plan against the reference's code for:
//   - is_prime      candidates/ordered-streaming/big_unsigned.h:83-97
//   - modulo        candidates/ordered-streaming/big_unsigned.h:54-72
//   - divide_word / modulo_word / times_word / multiply (same header)
//   - the prime walk of fherma_init (candidates/ordered-streaming/solve.cu:903-913),
//     restated here on top of those compiled helpers.
#include "big_unsigned.h"

#include <cstring>

using host_big::Big;

namespace {
Big load(const uint32_t* words, size_t count) {
    Big b(words, words + count);
    if (b.empty()) b.push_back(0);
    host_big::trim(b);
    return b;
}
int store(const Big& b, uint32_t* out, size_t out_words) {
    size_t used = b.size();
    while (used > 1 && !b[used - 1]) --used;
    if (used > out_words && !(used == 1 && b[0] == 0)) return -1;
    std::memset(out, 0, out_words * sizeof(uint32_t));
    for (size_t i = 0; i < used && i < out_words; ++i) out[i] = b[i];
    return 0;
}
}  // namespace

extern "C" {

int ref_is_prime(uint32_t p) { return host_big::is_prime(p) ? 1 : 0; }

int ref_modulo(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* out, size_t out_words) {
    return store(host_big::modulo(load(a, na), load(b, nb)), out, out_words);
}

uint32_t ref_modulo_word(const uint32_t* a, size_t na, uint32_t b) {
    return host_big::modulo_word(load(a, na), b);
}

int ref_divide_word(const uint32_t* a, size_t na, uint32_t b, uint32_t* out, size_t out_words) {
    return store(host_big::divide_word(load(a, na), b), out, out_words);
}

int ref_multiply(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* out, size_t out_words) {
    return store(host_big::multiply(load(a, na), load(b, nb)), out, out_words);
}

// Returns the number of primes written (<= capacity), -1 on
// "insufficient independent NTT fields", -2 on an invalid point or capacity.
int ref_prime_walk(uint32_t n, const uint32_t* q_words, size_t limbs, uint32_t* primes, size_t capacity) {
    using namespace host_big;
    if (!n || (n & (n - 1)) || n >= (1u << 29)) return -2;
    Big q = load(q_words, limbs);
    Big maximum = subtract(q, Big{1});
    Big bound = times_word(multiply(maximum, maximum), 4 * n), product{1};
    uint64_t step = 2 * uint64_t(n), candidate = ((uint64_t(1) << 31) - 2) / step * step + 1;
    size_t count = 0;
    while (compare(product, bound) <= 0) {
        if (candidate < (1u << 30)) return -1;
        if (is_prime(uint32_t(candidate))) {
            if (count >= capacity) return -2;
            primes[count++] = uint32_t(candidate);
            product = times_word(product, uint32_t(candidate));
        }
        candidate -= step;
    }
    return int(count);
}
}
