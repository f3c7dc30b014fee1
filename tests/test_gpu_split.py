"""Each base-case kernel of the split transform pinned in full against the
oracle at small sizes (pm_plan_create_split forces the split):

  n1 = 2, 4, 8      ntt_rows_2048<n1> (row length 2048)
  n1 = 16, 32, 2048 the generic ntt_rows_product base case at row length 2048
  n1 = 16 at N=1024 the generic kernel on a 1024-long row

For every prime, the product tap v_k must equal x * P_k^-1 mod p_k, x being the
exact signed negacyclic product (the oracle's convolution), and the output must
equal the GMP Kronecker product.  The plain configurations only reach n1 = 2
(N=32768, L=868) and n1 = 8 (N=65536, L=1747), the latter too large for a
per-prime comparison.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPLITS = [(4096, 8, 2), (4096, 8, 4), (4096, 8, 8), (4096, 8, 16), (4096, 8, 32), (4096, 8, 2048),
          (1024, 8, 16), (8192, 40, 8), (8192, 40, 64)]


@pytest.mark.parametrize("n,limbs,block", SPLITS)
def test_forced_split_is_exact(polymul, workloads, oracle, n, limbs, block):
    print("[gpu-case] split", n, limbs, block, flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(n * 31 + limbs * 7 + block)
    q = workloads.random_modulus(limbs, rng)
    a = workloads.uniform_below(q, n, rng)
    b = workloads.uniform_below(q, n, rng)
    with polymul.Plan(n, 32 * limbs, limbs, q, min_block=block) as plan:
        info = plan.info()
        assert info["block"] == block
        primes = [int(p) for p in plan.primes()]
        r = len(primes)
        x = oracle.signed_convolution(oracle.from_limbs(a), oracle.from_limbs(b))
        big_p = 1
        for p in primes:
            big_p *= p
        prod = plan.debug_stage(a, b, polymul.STAGE_PRODUCT).reshape(r, n)
        for k, p in enumerate(primes):
            inv = pow((big_p // p) % p, -1, p)
            expect = np.array([(xi % p) * inv % p for xi in x], dtype=np.uint32)
            bad = np.nonzero(prod[k] != expect)[0]
            assert bad.size == 0, "v_%d differs at %s" % (k, bad[:5])
        c = plan.multiply(a, b)
        expect = oracle.kronecker_gmp(a, b, q)
        bad = np.nonzero((c != expect).any(axis=1))[0]
        assert bad.size == 0, "coefficients differ, first at %s of %d" % (bad[:5], bad.size)


# Points whose host pipeline used to place a chunk past the last coefficient
# tile (6 chunks of 2 tiles over 8 tiles; 7 chunks of 3 over 16): the packing
# kernel then wrote past the packed-operand buffer.  Through pm_multiply and
# pm_multiply_batch, compared with the oracle.
@pytest.mark.parametrize("n,limbs", [(2048, 4096), (4096, 2500)])
def test_chunk_tiles_regression(polymul, workloads, oracle, n, limbs):
    print("[gpu-case] chunk regression", n, limbs, flush=True)
    rng = np.random.default_rng(n + limbs)
    q = workloads.random_modulus(limbs, rng)
    a = workloads.uniform_below(q, n, rng)
    b = workloads.uniform_below(q, n, rng)
    expect = oracle.kronecker_gmp(a, b, q)
    with polymul.Plan(n, 32 * limbs, limbs, q) as plan:
        assert np.array_equal(plan.multiply(a, b), expect)
    outs = polymul.multiply_batch([(n, limbs, q, a, b), (n, limbs, q, b, a)], 16 << 30, lanes=2)
    assert np.array_equal(outs[0], expect) and np.array_equal(outs[1], expect)
