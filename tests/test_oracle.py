"""Pins the CPU oracle (oracle/reference.py, oracle/kronecker.c) before it is
trusted as the parity reference.  Every check is independent of the
restatement it tests: Python's own big integers against GMP, the reference's
schoolbook against Kronecker packing, the reference's recorded prime lists and
its own compiled host helpers, sympy's primality test, and brute-force
modular identities.  One test plants a mistake in a copy of the oracle and
checks that the cross-check catches it."""
import json
import random
from pathlib import Path

import numpy as np
import pytest
import sympy

GOLDEN = Path(__file__).resolve().parent / "golden"


def _random_case(oracle, n, q, seed):
    rng = random.Random(seed)
    return [rng.randrange(q) for _ in range(n)], [rng.randrange(q) for _ in range(n)]


@pytest.mark.parametrize("n,width", [(1, 2), (2, 5), (8, 16), (64, 31), (64, 33), (16, 200)])
def test_kronecker_equals_schoolbook(oracle, n, width):
    """tests/reference.py:70-71: schoolbook cross-check for N <= 64."""
    q = oracle.modulus(n, width)
    pairs = oracle.edge_pairs(n, q) + [oracle.random_pair(n, q, s) for s in (101, 202)]
    for a, b in pairs:
        assert oracle.kronecker(a, b, q) == oracle.schoolbook(a, b, q)


@pytest.mark.parametrize("n,limbs", [(64, 2), (256, 8), (1024, 8), (128, 128)])
def test_gmp_kronecker_equals_python_kronecker(oracle, n, limbs):
    """Two independent big-integer libraries agree (GMP vs CPython ints)."""
    rng = random.Random(n * 1000 + limbs)
    q = rng.getrandbits(32 * limbs) | (1 << (32 * limbs - 1)) | 1
    a, b = _random_case(oracle, n, q, n + limbs)
    qw = oracle.int_to_words(q, limbs)
    got = oracle.kronecker_gmp(oracle.to_limbs(a, limbs), oracle.to_limbs(b, limbs), qw)
    assert oracle.from_limbs(got) == oracle.kronecker(a, b, q)


def test_gmp_schoolbook_and_single_coefficient(oracle):
    n, limbs = 32, 3
    q = (1 << 95) - 13
    a, b = _random_case(oracle, n, q, 5)
    qw = oracle.int_to_words(q, limbs)
    al, bl = oracle.to_limbs(a, limbs), oracle.to_limbs(b, limbs)
    expect = oracle.schoolbook(a, b, q)
    assert oracle.from_limbs(oracle.schoolbook_gmp(al, bl, qw)) == expect
    for i in (0, 7, n - 1):
        assert oracle.words_to_int(oracle.coefficient_gmp(al, bl, qw, i)) == expect[i]


def test_signed_convolution_matches_definition(oracle):
    n = 16
    rng = random.Random(3)
    a = [rng.randrange(1 << 70) for _ in range(n)]
    b = [rng.randrange(1 << 70) for _ in range(n)]
    x = oracle.signed_convolution(a, b)
    for i in range(n):
        direct = sum(a[j] * b[i - j] for j in range(i + 1)) - sum(a[j] * b[i - j + n] for j in range(i + 1, n))
        assert x[i] == direct


def test_planted_sign_mistake_is_caught(oracle):
    """A Kronecker fold with the negacyclic sign dropped (lo + hi) must fail the
    schoolbook cross-check: the check has teeth."""
    n, q = 8, oracle.modulus(8, 16)
    a, b = oracle.random_pair(n, q, 101)
    stride = oracle.kronecker_stride(n, q)
    left = int.from_bytes(b"".join(x.to_bytes(stride, "little") for x in a), "little")
    right = int.from_bytes(b"".join(x.to_bytes(stride, "little") for x in b), "little")
    raw = (left * right).to_bytes(2 * n * stride, "little")
    wrong = [(int.from_bytes(raw[i * stride:(i + 1) * stride], "little")
              + int.from_bytes(raw[(i + n) * stride:(i + n + 1) * stride], "little")) % q for i in range(n)]
    assert wrong != oracle.schoolbook(a, b, q)


def test_modulus_matches_fixture_and_definition(oracle):
    """tests/reference.py:9-16: largest prime q < 2^W with q = 1 (mod 2N)."""
    points = json.loads((GOLDEN / "points.json").read_text())
    for key, pt in points.items():
        n, width, q = pt["N"], pt["W"], int(pt["q"], 16)
        assert oracle.modulus(n, width) == q
        assert (q - 1) % (2 * n) == 0 and q < (1 << width)
        assert sympy.isprime(q)
        if width <= 512:  # no larger candidate below 2^W is prime
            step = 2 * n
            c = ((1 << width) - 2) // step * step + 1
            while c > q:
                assert not sympy.isprime(c)
                c -= step


def test_prime_walk_matches_reference_records(oracle):
    """The walk of solve.cu:905-913 reproduces the 57/114-prime lists the
    reference recorded (diagnostics/tensor-quotient/ARITHMETIC.json)."""
    recorded = json.loads((GOLDEN / "ref_primes.json").read_text())["reports"]
    points = json.loads((GOLDEN / "points.json").read_text())
    for rep in recorded:
        width = 1747 if rep["N"] == 65536 else 868
        q = int(points["%d/%d" % (rep["N"], width)]["q"], 16)
        primes = oracle.reference_primes(rep["N"], q)
        assert len(primes) == rep["fields"] and primes == rep["primes"]


def test_prime_walk_matches_compiled_reference_helpers(oracle):
    """Where oracle/_ref was built from the reference's own big_unsigned.h."""
    ref = oracle.reference_helpers()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent when building)")
    import ctypes
    points = json.loads((GOLDEN / "points.json").read_text())
    for key in ("1024/512", "8192/256", "32768/868", "65536/1747"):
        pt = points[key]
        q = int(pt["q"], 16)
        qw = oracle.int_to_words(q, pt["L"])
        out = np.zeros(4096, dtype=np.uint32)
        count = ref.ref_prime_walk(pt["N"], qw.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), pt["L"],
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), out.size)
        assert count > 0
        assert [int(p) for p in out[:count]] == oracle.reference_primes(pt["N"], q)
    # the reference walk runs out of fields for the wide random moduli
    wide = oracle.int_to_words((1 << (32 * 868)) - 1, 868)
    out = np.zeros(4096, dtype=np.uint32)
    assert ref.ref_prime_walk(32768, wide.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 868,
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), out.size) == -1


def test_is_prime_u32_against_sympy_and_reference(oracle):
    rng = random.Random(11)
    samples = [rng.randrange(1 << 30, 1 << 31) for _ in range(3000)]
    samples += [561, 1105, 1729, 2465, 2821, 6601, 8911, 3215031751, 2147483647, 2147483629, 4294967291]
    ref = oracle.reference_helpers()
    for x in samples:
        expect = sympy.isprime(x)
        assert oracle.is_prime_u32(x) == expect, x
        if ref is not None and x < (1 << 32):
            assert bool(ref.ref_is_prime(x)) == expect, x


@pytest.mark.parametrize("n,width", [(128, 128), (8192, 256), (32768, 868), (65536, 1747)])
def test_direct_crt_identities(oracle, n, width):
    """tests/direct_crt_math.py:9-31 restated: for signed x in [-B, B] the CRT
    quotient is the rounded sum of v_k/p_k, the fixed-point estimate is exact,
    and T = sum v_k M_k + K R is congruent to x mod q and fits L+2 words."""
    q = oracle.modulus(n, width)
    bound = n * (q - 1) ** 2
    primes = oracle.reference_primes(n, q)
    rng = random.Random(0xC47 + n + width)
    values = [-bound, -bound + 1, -q, -q + 1, -1, 0, 1, q - 1, q, bound - 1, bound]
    values += [rng.randrange(-bound, bound + 1) for _ in range(60)]
    product = 1
    for p in primes:
        product *= p
    for x in values:
        v, quotient, t = oracle.crt_reconstruct(x, primes, q)
        assert sum(a * (product // p) for a, p in zip(v, primes)) == x + quotient * product
        assert oracle.fixed_point_quotient(v, primes) == quotient
        # the exact 64-bit-reciprocal form used by the GPU (sum v_k floor(2^64/p_k))
        e = sum(a * (((1 << 64) - 1) // p) for a, p in zip(v, primes))
        assert (e + (1 << 63)) >> 64 == quotient
        assert t % q == x % q and t >= 0
        assert t.bit_length() <= 32 * ((width + 31) // 32 + 2)


def test_subtractive_montgomery_identities(oracle):
    """tests/tensor_montgomery_math.py:13-25 restated, including signed inputs
    (the GPU feeds signed byte sums)."""
    rng = random.Random(20261110)
    maximum = 7040 * 128 * 128 * 0x01010101
    for p in [(1 << 30) + 3, (1 << 31) - 1] + [sympy.nextprime(rng.randrange(1 << 30, 1 << 31)) for _ in range(30)]:
        r = 1 << 32
        for value in [0, 1, p - 1, p, p + 1, maximum, -maximum, p * r - 1] + [rng.randrange(-maximum, maximum)
                                                                             for _ in range(50)]:
            assert oracle.montgomery_redc(value, p) == value * pow(r, -1, p) % p


def test_small_golden_fixture(oracle):
    """The committed small-config fixture agrees with the oracle, and the
    unit/zero/impulse cases with their closed forms."""
    data = np.load(GOLDEN / "small_cases.npz")
    q = data["q"]
    qi = oracle.words_to_int(q)
    for name in ("random", "zero", "unit", "max", "impulse_top"):
        a, b, c = data[name + "_a"], data[name + "_b"], data[name + "_c"]
        assert np.array_equal(oracle.kronecker_gmp(a, b, q), c)
    assert not data["zero_c"].any()
    assert np.array_equal(data["unit_c"], data["unit_a"])
    # X^(N-1) * a: coefficient i+N-1 wraps to i-1 with a sign flip for i >= 1
    a = oracle.from_limbs(data["impulse_top_b"])
    c = oracle.from_limbs(data["impulse_top_c"])
    n = len(a)
    assert c[n - 1] == a[0] % qi
    assert all(c[i] == (-a[i + 1]) % qi for i in range(n - 1))
