// pm::prime_walk against the plain exact walk (solve.cu:905-913: descending
// candidates, exact product compared with the bound after every prime) on
// bounds where the walk's interval bounds must straddle and its exact fallback
// decides: exact products of the progression's own primes, and those +-1; plus
// bounds 4N(q-1)^2 of random q of many sizes.  Prints "ok <cases>" and exits 0
// when every walk agrees.
#include <cstdio>
#include <random>

#include "bigint.hpp"
#include "plan.hpp"

using pm::Words;

static bool plain_walk(uint64_t step, const Words& bound, std::vector<uint32_t>* primes) {
    primes->clear();
    Words product{1};
    uint64_t candidate = ((uint64_t(1) << 31) - 2) / step * step + 1;
    while (pm::compare(product, bound) <= 0) {
        if (candidate < (uint64_t(1) << 30)) return false;
        if (pm::is_prime_u32(uint32_t(candidate))) {
            primes->push_back(uint32_t(candidate));
            product = pm::mul_word(product, uint32_t(candidate));
        }
        candidate -= step;
    }
    return true;
}

static int cases = 0, failures = 0;

static void check(uint64_t step, const Words& bound, const char* what) {
    std::vector<uint32_t> expect, got;
    const bool ok_expect = plain_walk(step, bound, &expect);
    const bool ok_got = pm::prime_walk(step, bound, &got);
    ++cases;
    if (ok_expect != ok_got || expect != got) {
        ++failures;
        std::printf("FAIL step=%llu %s: plain %d/%zu primes, walk %d/%zu primes\n", (unsigned long long)step, what,
                    int(ok_expect), expect.size(), int(ok_got), got.size());
    }
}

int main() {
    for (uint64_t step : {uint64_t(2) * 1024, uint64_t(2) * 8192, uint64_t(2) * 65536}) {
        std::vector<uint32_t> first;
        plain_walk(step, pm::power_of_two(31 * 400), &first);  // the progression's leading primes
        for (size_t m : {size_t(1), size_t(2), size_t(17), size_t(100), size_t(257)}) {
            if (m > first.size()) continue;
            Words product{1};
            for (size_t i = 0; i < m; ++i) product = pm::mul_word(product, first[i]);
            check(step, pm::sub(product, Words{1}), "product - 1");
            check(step, product, "product");
            check(step, pm::add(product, Words{1}), "product + 1");
        }
    }
    std::mt19937 rng(2024);
    for (uint32_t limbs : {1u, 2u, 3u, 8u, 33u, 55u, 200u, 868u, 1747u}) {
        for (uint32_t n : {1024u, 65536u}) {
            Words q(limbs);
            for (auto& w : q) w = rng();
            q[limbs - 1] |= 0x80000000u;
            const Words qm1 = pm::sub(q, Words{1});
            const Words bound = pm::mul_word(pm::mul(qm1, qm1), 4 * n);
            for (uint64_t n1 = 1; n1 <= 8; n1 *= 2) check(2 * uint64_t(n / n1), bound, "4N(q-1)^2");
        }
    }
    if (failures) return 1;
    std::printf("ok %d\n", cases);
    return 0;
}
