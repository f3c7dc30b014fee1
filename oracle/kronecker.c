/*
 * TEST INFRASTRUCTURE ONLY. This is the CPU oracle for the exact negacyclic
 * product c = a*b mod (X^N + 1, q).  It is never linked into the product
 * library; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load it, and only as the checker.
 *
 * Restates the reference oracle tests/reference.py:29-43 (Kronecker packing:
 * every coefficient is written into a slot of `stride` bytes, with
 * stride = ceil(bitlen(N*(q-1)^2)/8), so one big-integer product holds every
 * acyclic coefficient exactly; the negacyclic fold is lo_i - hi_i, then mod q)
 * and the schoolbook cross-check tests/reference.py:19-26.
 *
 * The big-integer work goes through GMP (libgmp.so.10), exactly as the
 * reference's gmpy2 does.  This image ships the GMP shared library but not
 * gmp.h, so the handful of entry points used here are declared below from the
 * public GMP manual (mpz_t layout and the __gmpz_* exported names).
 *
 * Exported C ABI (all return 0 on success, a negative code on bad arguments):
 *   oracle_kronecker(N, L, q, a, b, c)          full product, REF:29-43
 *   oracle_coefficient(N, L, q, a, b, i, out)   one output coefficient by the
 *                                               direct signed sum, REF:19-26
 *   oracle_schoolbook(N, L, q, a, b, c)         all coefficients by REF:19-26
 *   oracle_signed_convolution(N, L, a, b, i, out, out_words, sign)
 *                                               the exact integer x_i before
 *                                               reduction (|x_i| <= N(q-1)^2)
 *   oracle_probab_prime(L, x, reps)             mpz_probab_prime_p, which is
 *                                               what gmpy2.is_prime(x, reps)
 *                                               calls (REF:12)
 * All multi-limb values are little-endian arrays of 32-bit limbs.
 */
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int alloc;
    int size;
    void *limbs;
} mpz_struct;
typedef mpz_struct mpz_t[1];

void __gmpz_init(mpz_t x);
void __gmpz_clear(mpz_t x);
void __gmpz_import(mpz_t rop, size_t count, int order, size_t size, int endian, size_t nails, const void *op);
void *__gmpz_export(void *rop, size_t *countp, int order, size_t size, int endian, size_t nails, const mpz_t op);
void __gmpz_mul(mpz_t rop, const mpz_t a, const mpz_t b);
void __gmpz_add(mpz_t rop, const mpz_t a, const mpz_t b);
void __gmpz_sub(mpz_t rop, const mpz_t a, const mpz_t b);
void __gmpz_addmul(mpz_t rop, const mpz_t a, const mpz_t b);
void __gmpz_submul(mpz_t rop, const mpz_t a, const mpz_t b);
void __gmpz_mod(mpz_t r, const mpz_t n, const mpz_t d);
void __gmpz_sub_ui(mpz_t rop, const mpz_t a, unsigned long b);
void __gmpz_set_ui(mpz_t rop, unsigned long v);
size_t __gmpz_sizeinbase(const mpz_t x, int base);
int __gmpz_probab_prime_p(const mpz_t n, int reps);

#define ORACLE_EINVAL (-1)
#define ORACLE_ENOMEM (-2)

static void load_words(mpz_t x, const uint32_t *words, size_t count) {
    __gmpz_import(x, count, -1, sizeof(uint32_t), 0, 0, words);
}

/* Writes |x| into out[0..words) (zero padded); returns -1 if it does not fit. */
static int store_words(uint32_t *out, size_t words, const mpz_t x) {
    size_t need = (__gmpz_sizeinbase(x, 2) + 31) / 32;
    if (x->size == 0) need = 0;
    if (need > words) return -1;
    memset(out, 0, words * sizeof(uint32_t));
    size_t got = 0;
    if (need) __gmpz_export(out, &got, -1, sizeof(uint32_t), 0, 0, x);
    return 0;
}

/* stride = ceil(bitlen(N*(q-1)^2) / 8) bytes, REF:36. */
static size_t kronecker_stride(uint32_t n, const mpz_t q) {
    mpz_t t, nn;
    __gmpz_init(t);
    __gmpz_init(nn);
    __gmpz_sub_ui(t, q, 1);
    __gmpz_mul(t, t, t);
    __gmpz_set_ui(nn, n);
    __gmpz_mul(t, t, nn);
    size_t bits = t->size == 0 ? 0 : __gmpz_sizeinbase(t, 2);
    __gmpz_clear(t);
    __gmpz_clear(nn);
    size_t stride = (bits + 7) / 8;
    return stride ? stride : 1;
}

/* Packs N coefficients of L limbs into one integer with `stride`-byte slots
 * (REF:37-38: the little-endian concatenation of x.to_bytes(stride)). */
static int pack(mpz_t out, const uint32_t *values, uint32_t n, uint32_t limbs, size_t stride) {
    size_t coef_bytes = (size_t)limbs * 4;
    size_t copy = coef_bytes < stride ? coef_bytes : stride;
    size_t total = (size_t)n * stride;
    uint8_t *buf = (uint8_t *)calloc(total ? total : 1, 1);
    if (!buf) return ORACLE_ENOMEM;
    for (size_t i = 0; i < n; ++i) {
        const uint8_t *src = (const uint8_t *)(values + i * (size_t)limbs);
        /* bytes beyond `copy` are zero for canonical inputs (value < q) */
        memcpy(buf + i * stride, src, copy);
    }
    __gmpz_import(out, total, -1, 1, 0, 0, buf);
    free(buf);
    return 0;
}

int oracle_kronecker(uint32_t n, uint32_t limbs, const uint32_t *q_words, const uint32_t *a,
                     const uint32_t *b, uint32_t *c) {
    if (!n || !limbs || !q_words || !a || !b || !c) return ORACLE_EINVAL;
    mpz_t q, left, right, prod, lo, hi;
    __gmpz_init(q);
    load_words(q, q_words, limbs);
    if (q->size == 0) {
        __gmpz_clear(q);
        return ORACLE_EINVAL;
    }
    size_t stride = kronecker_stride(n, q);
    __gmpz_init(left);
    __gmpz_init(right);
    __gmpz_init(prod);
    int rc = pack(left, a, n, limbs, stride);
    if (!rc) rc = pack(right, b, n, limbs, stride);
    if (rc) {
        __gmpz_clear(q); __gmpz_clear(left); __gmpz_clear(right); __gmpz_clear(prod);
        return rc;
    }
    __gmpz_mul(prod, left, right); /* REF:39 */
    __gmpz_clear(left);
    __gmpz_clear(right);
    /* raw = prod.to_bytes(2*n*stride); slot i minus slot i+n, mod q (REF:40-43). */
    size_t total = 2 * (size_t)n * stride;
    uint8_t *raw = (uint8_t *)calloc(total, 1);
    if (!raw) {
        __gmpz_clear(q); __gmpz_clear(prod);
        return ORACLE_ENOMEM;
    }
    size_t got = 0;
    if (prod->size) __gmpz_export(raw, &got, -1, 1, 0, 0, prod);
    __gmpz_clear(prod);
    __gmpz_init(lo);
    __gmpz_init(hi);
    for (size_t i = 0; i < n; ++i) {
        __gmpz_import(lo, stride, -1, 1, 0, 0, raw + i * stride);
        __gmpz_import(hi, stride, -1, 1, 0, 0, raw + (i + n) * stride);
        __gmpz_sub(lo, lo, hi);
        __gmpz_mod(lo, lo, q); /* Python % q: non-negative result */
        if (store_words(c + i * (size_t)limbs, limbs, lo)) { rc = ORACLE_EINVAL; break; }
    }
    free(raw);
    __gmpz_clear(lo);
    __gmpz_clear(hi);
    __gmpz_clear(q);
    return rc;
}

/* x_i = sum_{j<=i} a_j b_{i-j} - sum_{j>i} a_j b_{i-j+n}  (REF:22-25). */
static void signed_sum(mpz_t acc, uint32_t n, uint32_t limbs, const uint32_t *a, const uint32_t *b,
                       uint32_t i) {
    mpz_t x, y;
    __gmpz_init(x);
    __gmpz_init(y);
    __gmpz_set_ui(acc, 0);
    for (uint32_t j = 0; j < n; ++j) {
        uint32_t k = (i >= j) ? i - j : i + n - j;
        load_words(x, a + (size_t)j * limbs, limbs);
        load_words(y, b + (size_t)k * limbs, limbs);
        if (j <= i) __gmpz_addmul(acc, x, y);
        else __gmpz_submul(acc, x, y);
    }
    __gmpz_clear(x);
    __gmpz_clear(y);
}

int oracle_coefficient(uint32_t n, uint32_t limbs, const uint32_t *q_words, const uint32_t *a,
                       const uint32_t *b, uint32_t i, uint32_t *out) {
    if (!n || !limbs || i >= n || !q_words || !a || !b || !out) return ORACLE_EINVAL;
    mpz_t q, acc;
    __gmpz_init(q);
    __gmpz_init(acc);
    load_words(q, q_words, limbs);
    if (q->size == 0) {
        __gmpz_clear(q); __gmpz_clear(acc);
        return ORACLE_EINVAL;
    }
    signed_sum(acc, n, limbs, a, b, i);
    __gmpz_mod(acc, acc, q);
    int rc = store_words(out, limbs, acc) ? ORACLE_EINVAL : 0;
    __gmpz_clear(q);
    __gmpz_clear(acc);
    return rc;
}

int oracle_schoolbook(uint32_t n, uint32_t limbs, const uint32_t *q_words, const uint32_t *a,
                      const uint32_t *b, uint32_t *c) {
    for (uint32_t i = 0; i < n; ++i) {
        int rc = oracle_coefficient(n, limbs, q_words, a, b, i, c + (size_t)i * limbs);
        if (rc) return rc;
    }
    return 0;
}

/* The exact signed convolution coefficient: |x_i| in out[0..out_words), sign in *sign
 * (-1, 0 or +1).  Used to pin the per-prime residues x mod p of the GPU path. */
int oracle_signed_convolution(uint32_t n, uint32_t limbs, const uint32_t *a, const uint32_t *b,
                              uint32_t i, uint32_t *out, uint32_t out_words, int32_t *sign) {
    if (!n || !limbs || i >= n || !a || !b || !out || !sign) return ORACLE_EINVAL;
    mpz_t acc;
    __gmpz_init(acc);
    signed_sum(acc, n, limbs, a, b, i);
    *sign = acc->size > 0 ? 1 : (acc->size < 0 ? -1 : 0);
    int rc = store_words(out, out_words, acc) ? ORACLE_EINVAL : 0;
    __gmpz_clear(acc);
    return rc;
}

int oracle_probab_prime(uint32_t limbs, const uint32_t *x_words, int reps) {
    if (!limbs || !x_words) return ORACLE_EINVAL;
    mpz_t x;
    __gmpz_init(x);
    load_words(x, x_words, limbs);
    int r = __gmpz_probab_prime_p(x, reps);
    __gmpz_clear(x);
    return r;
}
