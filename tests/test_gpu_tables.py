"""Plans made for the GPU build their prime-set tables (twiddles, conversion
weights and offsets, beta) and the q-dependent CRT weight matrix with kernels;
host-only plans build the same tables in plan.cpp, whose closed forms the CPU
tests pin.  The two must agree byte for byte, for the reference's own walk,
for the split transform (n1 > 1), for tiny N and for a q that one of the
primes divides."""
import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
TABLES = {  # id -> dtype of the exported table
    1: np.uint32, 2: np.int8, 3: np.uint32, 4: np.uint32, 5: np.int8, 6: np.uint32,
    7: np.uint32, 8: np.uint32, 9: np.uint32, 10: np.uint64, 11: np.uint32, 12: np.uint32,
}


def _cases(workloads):
    rng = np.random.default_rng(0x7AB1E5)
    points = json.loads((ROOT / "tests" / "golden" / "points.json").read_text())
    ref = points["65536/1747"]
    cases = [
        ("ref main q (reference walk)", ref["N"], ref["L"], workloads.int_to_words(int(ref["q"], 16), ref["L"])),
        ("B-small", 1024, 8, workloads.random_modulus(8, rng)),
        ("B-mid (split n1=2)", 32768, 868, workloads.random_modulus(868, rng)),
        ("tiny N=2", 2, 3, workloads.random_modulus(3, rng)),
        ("N=64 odd width", 64, 5, workloads.random_modulus(5, rng)),
    ]
    return cases


@pytest.mark.parametrize("index", range(5))
def test_device_tables_equal_host_tables(polymul, workloads, index):
    name, n, limbs, q = _cases(workloads)[index]
    print("[gpu-case] tables", name, flush=True)
    with polymul.Plan(n, 32 * limbs, limbs, q) as dev, polymul.Plan(n, 32 * limbs, limbs, q, device=False) as host:
        for table_id, dtype in TABLES.items():
            got, expect = dev.table(table_id, dtype), host.table(table_id, dtype)
            assert got.shape == expect.shape, (name, table_id)
            assert np.array_equal(got, expect), (name, table_id)


def test_device_tables_when_a_prime_divides_q(polymul, workloads):
    """q = p_0 * (odd cofactor): M_0 takes the direct-product branch."""
    print("[gpu-case] tables p | q", flush=True)
    n, limbs = 1024, 8
    probe = workloads.random_modulus(limbs, np.random.default_rng(5))
    with polymul.Plan(n, 32 * limbs, limbs, probe, device=False) as plan:
        p0 = int(plan.primes()[0])
    cofactor = workloads.words_to_int(workloads.random_modulus(limbs - 1, np.random.default_rng(6)))
    q = workloads.int_to_words(p0 * cofactor, limbs)
    with polymul.Plan(n, 32 * limbs, limbs, q) as dev, polymul.Plan(n, 32 * limbs, limbs, q, device=False) as host:
        assert int(host.primes()[0]) == p0
        for table_id, dtype in TABLES.items():
            assert np.array_equal(dev.table(table_id, dtype), host.table(table_id, dtype)), table_id
