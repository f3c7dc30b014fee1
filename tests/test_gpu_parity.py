"""GPU parity of the gfx950 path against the CPU oracle (bit-exact).

Every product goes through the C ABI of libpolymul_amd.so; the oracle is
oracle/reference.py (GMP Kronecker, tests/reference.py:29-43 restated) and,
per stage, the closed forms the oracle pins:
  residues   a mod p_k, beta_k*b mod p_k              (conversion GEMM)
  product    x_i * P_k^-1 mod p_k, x = exact signed   (NTTs + base case)
             negacyclic integer product
  words      sum_j 2^32j D_j + corr == sum_k v_k M_k   (CRT GEMM, exact)
  output     c == Kronecker oracle
"""
import random
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _announce(*what):
    print("[gpu-case]", *what, flush=True)
    sys.stdout.flush()


def _limbs(values, limbs, oracle):
    return oracle.to_limbs(values, limbs)


def _check_product(plan, a, b, q_words, oracle):
    c = plan.multiply(a, b)
    expect = oracle.kronecker_gmp(a, b, q_words)
    mismatch = np.nonzero((c != expect).any(axis=1))[0]
    assert mismatch.size == 0, "coefficients differ, first at %s of %d" % (mismatch[:5], mismatch.size)


# ---------------------------------------------------------------------------
# per-stage parity at small sizes
# ---------------------------------------------------------------------------
STAGE_POINTS = [
    (8, 16, "ntt"),        # tiny
    (64, 33, "ntt"),       # two limbs
    (1024, 256, "random"),  # the small plain config
]


@pytest.mark.parametrize("n,width,kind", STAGE_POINTS)
def test_stages(polymul, workloads, oracle, n, width, kind):
    _announce("stages", n, width, kind)
    limbs = (width + 31) // 32
    if kind == "ntt":
        q = oracle.modulus(n, width)
    else:
        rng = np.random.Generator(np.random.PCG64(7))
        q = oracle.words_to_int(workloads.random_modulus(limbs, rng))
    q_words = oracle.int_to_words(q, limbs)
    a_list, b_list = oracle.random_pair(n, q, 101)
    a, b = oracle.to_limbs(a_list, limbs), oracle.to_limbs(b_list, limbs)
    with polymul.Plan(n, width, limbs, q_words) as plan:
        primes = [int(p) for p in plan.primes()]
        r = len(primes)
        beta = [int(x) for x in plan.table(polymul.TABLE_BETA, np.uint32)]
        # 1. conversion
        res = plan.debug_stage(a, b, polymul.STAGE_RESIDUES).reshape(2, r, n)
        for k, p in enumerate(primes):
            assert [int(v) for v in res[0, k]] == [x % p for x in a_list], "a mod p_%d" % k
            assert [int(v) for v in res[1, k]] == [beta[k] * x % p for x in b_list], "beta*b mod p_%d" % k
        # 2. transforms + base case + inverse
        x = oracle.signed_convolution(a_list, b_list)
        big_p = 1
        for p in primes:
            big_p *= p
        prod = plan.debug_stage(a, b, polymul.STAGE_PRODUCT).reshape(r, n)
        for k, p in enumerate(primes):
            inv = pow((big_p // p) % p, -1, p)
            assert [int(v) for v in prod[k]] == [(xi % p) * inv % p for xi in x], "v_%d" % k
        # 3. CRT word sums: exact identities of the signed-byte GEMM
        words = plan.debug_stage(a, b, polymul.STAGE_WORDS).reshape(limbs + 4, n)
        m_mod_q = plan.table(polymul.TABLE_M_MOD_Q, np.uint32).reshape(r, limbs)
        m_vals = [oracle.words_to_int(m_mod_q[k]) for k in range(r)]
        rcp = [int(v) for v in plan.table(polymul.TABLE_RECIPROCALS, np.uint64)]
        corr_t = oracle.words_to_int(plan.table(polymul.TABLE_CORR_T, np.uint32))
        corr_e = oracle.words_to_int(plan.table(polymul.TABLE_CORR_E, np.uint32))
        for i in range(0, n, max(1, n // 16)):
            v = [int(prod[k, i]) for k in range(r)]
            t_sum = sum(int(words[j, i]) << (32 * j) for j in range(limbs + 1)) + corr_t
            assert t_sum == sum(vk * mk for vk, mk in zip(v, m_vals)), "T sum at %d" % i
            e_sum = sum(int(words[limbs + 1 + j, i]) << (32 * j) for j in range(3)) + corr_e
            assert e_sum == sum(vk * rk for vk, rk in zip(v, rcp)), "E sum at %d" % i
            _, quotient, _ = oracle.crt_reconstruct(x[i], primes, q)
            assert (e_sum + (1 << 63)) >> 64 == quotient
        # 4. output
        c = plan.debug_stage(a, b, polymul.STAGE_OUTPUT).reshape(n, limbs)
        assert oracle.from_limbs(c) == [xi % q for xi in x]


# ---------------------------------------------------------------------------
# the reference's validation matrix (tools/validate_local.py:13-14), NTT-prime q,
# edge cases of tests/reference.py:52-57 and seeds 101, 202
# ---------------------------------------------------------------------------
VALIDATE_POINTS = [(1, 2), (2, 5), (8, 16), (64, 31), (64, 33), (128, 128), (256, 256),
                   (1024, 512), (8192, 256), (32768, 868), (65536, 1747)]


@pytest.mark.parametrize("n,width", VALIDATE_POINTS)
def test_reference_matrix(polymul, oracle, n, width):
    _announce("reference-matrix", n, width)
    limbs = (width + 31) // 32
    q = oracle.modulus(n, width)
    q_words = oracle.int_to_words(q, limbs)
    pairs = oracle.edge_pairs(n, q) + [oracle.random_pair(n, q, s) for s in (101, 202)]
    with polymul.Plan(n, width, limbs, q_words) as plan:
        for a_list, b_list in pairs:
            a, b = oracle.to_limbs(a_list, limbs), oracle.to_limbs(b_list, limbs)
            _check_product(plan, a, b, q_words, oracle)


# tools/stress_local.py:16-27 boundary pairs (wider coefficients, larger transform)
STRESS_POINTS = [(128, 4096), (131072, 128), (32768, 868), (65536, 1747)]


@pytest.mark.parametrize("n,width", STRESS_POINTS)
def test_stress_boundaries(polymul, oracle, n, width):
    _announce("stress", n, width)
    limbs = (width + 31) // 32
    q = oracle.modulus(n, width)
    q_words = oracle.int_to_words(q, limbs)
    pairs = oracle.stress_pairs(n, q, limbs) + [oracle.random_pair(n, q, 10001)]
    with polymul.Plan(n, width, limbs, q_words) as plan:
        for a_list, b_list in pairs:
            a, b = oracle.to_limbs(a_list, limbs), oracle.to_limbs(b_list, limbs)
            _check_product(plan, a, b, q_words, oracle)


# ---------------------------------------------------------------------------
# the largest transforms: N = 2^20 (64-bit q) and N = 2^24, the largest N a plan
# accepts (32-bit q); random odd q, full comparison with the oracle
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,limbs", [(1 << 20, 2), (1 << 24, 1)])
def test_largest_transforms(polymul, workloads, oracle, n, limbs):
    _announce("largest-transforms", n, limbs)
    rng = np.random.default_rng(7000 + limbs)
    q = workloads.random_modulus(limbs, rng)
    a = workloads.uniform_below(q, n, rng)
    b = workloads.uniform_below(q, n, rng)
    with polymul.Plan(n, 32 * limbs, limbs, q) as plan:
        _check_product(plan, a, b, q, oracle)


# the widest q: 131,072 bits (L = 4096, about 8,460 primes); N = 1 is a plain
# modular product of two such integers
@pytest.mark.parametrize("n", [1, 64])
def test_widest_moduli(polymul, workloads, oracle, n):
    limbs = 4096
    _announce("widest-moduli", n, limbs)
    rng = np.random.default_rng(8000 + n)
    q = workloads.random_modulus(limbs, rng)
    a = workloads.uniform_below(q, n, rng)
    b = workloads.uniform_below(q, n, rng)
    with polymul.Plan(n, 32 * limbs, limbs, q) as plan:
        assert plan.info()["primes"] > 8000
        _check_product(plan, a, b, q, oracle)
        m = np.tile(workloads.q_minus_one(q), (n, 1))
        _check_product(plan, m, m, q, oracle)


# ---------------------------------------------------------------------------
# the plain configurations with random odd q
# ---------------------------------------------------------------------------
def _edge_operands(workloads, q_words, n, kind):
    limbs = q_words.size
    z = np.zeros((n, limbs), dtype=np.uint32)
    if kind == "zero":
        return z
    if kind == "unit":
        z[0, 0] = 1
        return z
    if kind == "max":
        return np.tile(workloads.q_minus_one(q_words), (n, 1))
    if kind.startswith("x^"):
        z[int(kind[2:]), 0] = 1
        return z
    raise ValueError(kind)


def test_small_random_modulus(polymul, workloads, oracle):
    _announce("small")
    n, width, limbs, q, a, b = workloads.seeded_case("small")
    with polymul.Plan(n, width, limbs, q) as plan:
        assert plan.info()["primes"] == 17
        _check_product(plan, a, b, q, oracle)
        for kind in ("zero", "unit", "max", "x^0", "x^1", "x^%d" % (n - 1)):
            e = _edge_operands(workloads, q, n, kind)
            _check_product(plan, e, b, q, oracle)
            _check_product(plan, a, e, q, oracle)
        m = _edge_operands(workloads, q, n, "max")
        _check_product(plan, m, m, q, oracle)


def test_mid_random_modulus(polymul, workloads, oracle):
    _announce("mid")
    n, width, limbs, q, a, b = workloads.seeded_case("mid")
    with polymul.Plan(n, width, limbs, q) as plan:
        info = plan.info()
        assert info["block"] > 1  # the split transform is needed here
        _check_product(plan, a, b, q, oracle)


def test_large_random_modulus_properties(polymul, workloads, oracle):
    """The large config at full size: unit and shift identities, commutativity,
    and sampled coefficients against the oracle's direct signed sum."""
    _announce("large")
    n, width, limbs, q, a, b = workloads.seeded_case("large")
    with polymul.Plan(n, width, limbs, q) as plan:
        c = plan.multiply(a, b)
        for i in (0, 1, n // 2, n - 1):
            _announce("large coefficient", i)
            assert np.array_equal(c[i], oracle.coefficient_gmp(a, b, q, i)), "coefficient %d" % i
        unit = _edge_operands(workloads, q, n, "unit")
        assert np.array_equal(plan.multiply(a, unit), a)
        # X * a: coefficients shift up by one, the top one wraps with a sign flip
        x1 = _edge_operands(workloads, q, n, "x^1")
        shifted = plan.multiply(x1, a)
        assert np.array_equal(shifted[1:], a[:-1])
        top = oracle.words_to_int(a[-1])
        assert oracle.words_to_int(shifted[0]) == (-top) % oracle.words_to_int(q)
        assert np.array_equal(plan.multiply(b, a), c)


# q wider than 62 limbs takes the wave-per-coefficient reconstruction: edge
# operands whose results sit next to a multiple of q (all q-1: every result is
# a small integer or q minus one) take its correction pass; a 62-bit q stored
# in 70 limbs starts its Barrett window at word 0.
@pytest.mark.parametrize("n,limbs,kind", [(256, 100, "random"), (64, 130, "random"), (128, 70, "narrow")])
def test_wide_q_reconstruction_edges(polymul, workloads, oracle, n, limbs, kind):
    _announce("wide-q edges", n, limbs, kind)
    rng = np.random.default_rng(n * 1000 + limbs)
    if kind == "narrow":
        q_int = (1 << 61) + 1235
        q = oracle.int_to_words(q_int, limbs)
        draw = random.Random(limbs)
        a = oracle.to_limbs([draw.randrange(q_int) for _ in range(n)], limbs)
        b = oracle.to_limbs([draw.randrange(q_int) for _ in range(n)], limbs)
    else:
        q = workloads.random_modulus(limbs, rng)
        a = workloads.uniform_below(q, n, rng)
        b = workloads.uniform_below(q, n, rng)
    with polymul.Plan(n, 32 * limbs, limbs, q) as plan:
        _check_product(plan, a, b, q, oracle)
        for kind_a in ("zero", "unit", "max", "x^%d" % (n - 1)):
            e = _edge_operands(workloads, q, n, kind_a)
            _check_product(plan, e, b, q, oracle)
        m = _edge_operands(workloads, q, n, "max")
        _check_product(plan, m, m, q, oracle)


# q needs no structure: an even q and a power of two (q = 2^64 in three limbs,
# so Barrett's window and the final subtraction see a q with zero low words)
# give the oracle's product, also for the all-(q-1) operands whose results sit
# next to a multiple of q; 70 limbs take the wave-per-coefficient reconstruction.
@pytest.mark.parametrize("n,limbs,kind", [(64, 3, "power"), (1024, 8, "even"), (256, 70, "even")])
def test_even_modulus(polymul, workloads, oracle, n, limbs, kind):
    _announce("even q", n, limbs, kind)
    rng = np.random.default_rng(7700 + n + limbs)
    if kind == "power":
        q = np.zeros(limbs, dtype=np.uint32)
        q[-1] = 1
    else:
        q = workloads.random_modulus(limbs, rng)
        q[0] &= np.uint32(0xFFFFFFFE)
    assert oracle.words_to_int(q) % 2 == 0
    a = workloads.uniform_below(q, n, rng)
    b = workloads.uniform_below(q, n, rng)
    with polymul.Plan(n, 32 * limbs, limbs, q) as plan:
        _check_product(plan, a, b, q, oracle)
        m = _edge_operands(workloads, q, n, "max")
        _check_product(plan, m, m, q, oracle)
        _check_product(plan, m, b, q, oracle)


# Host products large enough for one chunk per column strip move their
# coefficients in strips (256 columns of the [rows][2048] matrix the transforms
# use): 2 rows (one chunk per strip), 4 and 8 rows (the first and last strips
# in pieces of rows).  The chunk-major staging, the per-strip column stages and
# the scatter into c must give the device entry's result, which holds the
# operands in their natural order; sampled coefficients against the oracle.
@pytest.mark.parametrize("n,limbs", [(4096, 2700), (8192, 1350), (16384, 700)])
def test_column_strip_chunks(polymul, workloads, oracle, device_memory, n, limbs):
    _announce("column strips", n, limbs)
    rng = np.random.default_rng(n + limbs)
    q = workloads.random_modulus(limbs, rng)
    a = workloads.uniform_below(q, n, rng)
    b = workloads.uniform_below(q, n, rng)
    with polymul.Plan(n, 32 * limbs, limbs, q) as plan:
        c = plan.multiply(a, b)
        da, db = device_memory.upload(a), device_memory.upload(b)
        dc = device_memory.alloc(a.nbytes)
        try:
            plan.multiply_device(da, db, dc)
            assert np.array_equal(device_memory.download(dc, a.shape, np.uint32), c)
        finally:
            for ptr in (da, db, dc):
                device_memory.free(ptr)
        for i in (0, 255, 256, 2047, 2048, n // 2 + 300, n - 1):
            assert np.array_equal(c[i], oracle.coefficient_gmp(a, b, q, i)), "coefficient %d" % i
        assert np.array_equal(plan.multiply(b, a), c)
