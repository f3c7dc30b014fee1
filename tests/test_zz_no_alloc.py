"""pm_multiply_device_scratch: the device product whose working buffers come
from the library's static per-GPU workspace instead of the plan's own, so the
call allocates nothing.  Its output must equal pm_multiply_device's, word for
word, on consecutive calls that reuse the workspace with different shapes
(the second smaller than the first), and a plan whose workspace exceeds the
static array must be refused with PM_ERR_INVALID_ARGUMENT.

Shapes: the reference's competition point (N = 65536, 1747-bit NTT-prime q,
L = 55; about 125.5 MiB of workspace, just inside the 128 MiB array), then the
small configuration (N = 1024, 256-bit random odd q), whose product is also
compared with the oracle, then the competition point once more.
"""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

POINTS = json.loads((Path(__file__).resolve().parent / "golden" / "points.json").read_text())


def _announce(*what):
    print("[gpu-case]", *what, flush=True)
    sys.stdout.flush()


def _both_entries(plan, device_memory, a, b):
    """(pm_multiply_device's c, pm_multiply_device_scratch's c) on the same
    device operands; c buffers start from different garbage."""
    pa, pb = device_memory.upload(a), device_memory.upload(b)
    pc_plain = device_memory.upload(np.full(a.shape, 0xA5A5A5A5, dtype=np.uint32))
    pc_scratch = device_memory.upload(np.full(a.shape, 0x5A5A5A5A, dtype=np.uint32))
    try:
        plan.multiply_device(pa, pb, pc_plain)
        plan.multiply_device_scratch(pa, pb, pc_scratch)
        plain = device_memory.download(pc_plain, a.shape, np.uint32)
        scratch = device_memory.download(pc_scratch, a.shape, np.uint32)
    finally:
        for ptr in (pa, pb, pc_plain, pc_scratch):
            device_memory.free(ptr)
    return plain, scratch


def test_scratch_entry_matches_device_entry(polymul, workloads, device_memory, oracle):
    ref = POINTS["65536/1747"]
    ref_q = workloads.int_to_words(int(ref["q"], 16), ref["L"])
    a1, b1 = workloads.operands_for(ref["N"], ref_q, 0x5C7A)
    _announce("scratch entry, competition point", ref["N"], ref["L"])
    with polymul.Plan(ref["N"], ref["W"], ref["L"], ref_q) as big:
        plain, scratch = _both_entries(big, device_memory, a1, b1)
        assert np.array_equal(plain, scratch), "first call (N=65536, L=55) differs from pm_multiply_device"

        _announce("scratch entry, small configuration after the large one")
        n, width, limbs, q, a2, b2 = workloads.seeded_case("small")
        with polymul.Plan(n, width, limbs, q) as small:
            plain2, scratch2 = _both_entries(small, device_memory, a2, b2)
        assert np.array_equal(plain2, scratch2), "second call (N=1024, L=8) differs from pm_multiply_device"
        assert np.array_equal(scratch2, oracle.kronecker_gmp(a2, b2, q)), "second call differs from the oracle"

        _announce("scratch entry, competition point again with new operands")
        a3, b3 = workloads.operands_for(ref["N"], ref_q, 0x5C7B)
        plain3, scratch3 = _both_entries(big, device_memory, a3, b3)
        assert np.array_equal(plain3, scratch3), "third call (N=65536, L=55) differs from pm_multiply_device"


def test_scratch_entry_refuses_oversized_plans(polymul, workloads, device_memory):
    rng = np.random.Generator(np.random.PCG64(0x5C7C))
    limbs = 72  # 176 MiB of workspace at N = 65536
    q = workloads.random_modulus(limbs, rng)
    _announce("scratch entry refusal", 65536, limbs)
    with polymul.Plan(65536, 32 * limbs, limbs, q) as plan:
        words = 65536 * limbs
        pa, pb, pc = device_memory.alloc(4 * words), device_memory.alloc(4 * words), device_memory.alloc(4 * words)
        try:
            with pytest.raises(polymul.PolymulError) as err:
                plan.multiply_device_scratch(pa, pb, pc)
        finally:
            for ptr in (pa, pb, pc):
                device_memory.free(ptr)
    assert err.value.status == polymul.PM_ERR_INVALID_ARGUMENT
    assert "static scratch" in str(err.value)
