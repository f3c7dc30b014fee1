"""The large workloads at their full sizes, every output word checked: the GPU
result's SHA-256 must equal the digest of the oracle's GMP Kronecker product
(tests/reference.py:29-43 restated in oracle/kronecker.c), computed once by
tests/golden/make_workload_golden.py and committed in
tests/golden/workload_digests.json.

  plain large      N=65536, L=1747 (55,904-bit random odd q), through the
                   host-vector entry and the device entry
  batched          the 64 products of the batched workload through
                   pm_multiply_batch (60 GB budget, 4 lanes)
  small batched    64 small products (N <= 8192, L <= 128) through
                   pm_multiply_batch, through pm_multiply_device from four
                   threads at once and through pm_multiply_device_batch
  resident         64 random a times one fixed b, b's residues on the device
  served           the 48 requests of the served workload with Poisson
                   arrivals, 200 GB budget: each result in the mix must be the
                   oracle's, i.e. the same as alone
  stress seeds     the reference's fresh-seed stress cases
                   (tools/stress_local.py:16-21): seeds 10001-10020 at
                   (32768, 868) and (65536, 1747) bits, 10001-10004 at
                   (128, 4096) and (131072, 128)

Input digests are compared first, so that a generator difference is reported
as such rather than as a wrong product.
"""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = json.loads((Path(__file__).resolve().parent / "golden" / "workload_digests.json").read_text())
GiB = 1 << 30


def _announce(*what):
    print("[gpu-case]", *what, flush=True)
    sys.stdout.flush()


def _check_inputs(workloads, entry, q, a, b, what):
    got = (workloads.digest(q), workloads.digest(a), workloads.digest(b))
    assert got == (entry["q"], entry["a"], entry["b"]), "%s: the seeded inputs differ from the golden run's" % what


def test_large_full_output(polymul, workloads, device_memory):
    g = GOLDEN["large"]
    _announce("large full digest", g["n"], g["limbs"])
    n, width, limbs, q, a, b = workloads.seeded_case("large")
    _check_inputs(workloads, g, q, a, b, "large")
    with polymul.Plan(n, width, limbs, q) as plan:
        c = plan.multiply(a, b)
        if workloads.digest(c) != g["c"]:
            blk = g["block"]
            bad = [i for i, d in enumerate(g["blocks"]) if workloads.digest(c[i * blk:(i + 1) * blk]) != d]
            raise AssertionError("host entry: %d of %d coefficient blocks differ, first %s"
                                 % (len(bad), len(g["blocks"]), bad[:5]))
        del c
        pa, pb, pc = device_memory.upload(a), device_memory.upload(b), device_memory.alloc(a.nbytes)
        try:
            plan.multiply_device(pa, pb, pc)
            cd = device_memory.download(pc, a.shape, np.uint32)
        finally:
            for ptr in (pa, pb, pc):
                device_memory.free(ptr)
        assert workloads.digest(cd) == g["c"], "device entry differs from the oracle"


def test_batched_workload(polymul, workloads):
    g = GOLDEN["batch"]["products"]
    products = workloads.batch_products()
    assert len(products) == len(g)
    _announce("batched workload", len(products))
    for (n, limbs, q, a, b), e in zip(products, g):
        _check_inputs(workloads, e, q, a, b, "batch product N=%d L=%d" % (n, limbs))
    outs = polymul.multiply_batch(products, 60 * GiB, lanes=4)
    bad = [i for i, (c, e) in enumerate(zip(outs, g)) if workloads.digest(c) != e["c"]]
    assert not bad, "batched products differ from the oracle: %s" % [(i, g[i]["n"], g[i]["limbs"]) for i in bad]


def test_small_batched_workload(polymul, workloads, device_memory):
    """The mix of 64 small products (N <= 8192, L <= 128) through
    pm_multiply_batch, through pm_multiply_device from four threads at once
    (each plan on its own stream), and through pm_multiply_device_batch (each
    stage one launch for all 64), against the oracle's digests."""
    import threading

    g = GOLDEN["batch_small"]["products"]
    products = workloads.batch_small_products()
    assert len(products) == len(g)
    _announce("small batched workload", len(products))
    for (n, limbs, q, a, b), e in zip(products, g):
        _check_inputs(workloads, e, q, a, b, "small batch product N=%d L=%d" % (n, limbs))
    outs = polymul.multiply_batch(products, 4 * GiB, lanes=4)
    bad = [i for i, (c, e) in enumerate(zip(outs, g)) if workloads.digest(c) != e["c"]]
    assert not bad, "batched products differ from the oracle: %s" % bad
    plans = [polymul.Plan(n, 32 * limbs, limbs, q) for n, limbs, q, _a, _b in products]
    bufs = [(device_memory.upload(a), device_memory.upload(b), device_memory.alloc(a.nbytes))
            for _n, _l, _q, a, b in products]
    try:
        pending = iter(range(len(products)))
        lock = threading.Lock()
        errors = []

        def worker():
            while True:
                with lock:
                    i = next(pending, None)
                if i is None:
                    return
                try:
                    plans[i].multiply_device(*bufs[i])
                except Exception as e:  # reported below, in the test's thread
                    errors.append((i, e))

        threads = [threading.Thread(target=worker) for _ in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        assert not errors, errors
        bad = [i for i, ((n, limbs, *_r), (_a, _b, pc), e) in enumerate(zip(products, bufs, g))
               if workloads.digest(device_memory.download(pc, (n, limbs), np.uint32)) != e["c"]]
        assert not bad, "device products from four threads differ from the oracle: %s" % bad
        # the grouped entry: each stage one launch for all 64 products, twice in a row
        for attempt in range(2):
            for (n, limbs, *_r), (_a, _b, pc) in zip(products, bufs):
                device_memory.write(pc, np.zeros((n, limbs), dtype=np.uint32))
            polymul.multiply_device_batch(plans, [x[0] for x in bufs], [x[1] for x in bufs], [x[2] for x in bufs])
            bad = [i for i, ((n, limbs, *_r), (_a, _b, pc), e) in enumerate(zip(products, bufs, g))
                   if workloads.digest(device_memory.download(pc, (n, limbs), np.uint32)) != e["c"]]
            assert not bad, "grouped products (pass %d) differ from the oracle: %s" % (attempt, bad)
    finally:
        for pa, pb, pc in bufs:
            for ptr in (pa, pb, pc):
                device_memory.free(ptr)
        for plan in plans:
            plan.close()


def test_resident_operand_workload(polymul, workloads):
    g = GOLDEN["resident"]["products"]
    n, limbs, q, b, a_list = workloads.resident_case()
    _announce("resident workload", n, limbs, len(a_list))
    for a, e in zip(a_list, g):
        _check_inputs(workloads, e, q, a, b, "resident product")
    with polymul.Plan(n, 32 * limbs, limbs, q) as plan:
        with plan.operand(b) as fixed:
            bad = [i for i, (a, e) in enumerate(zip(a_list, g))
                   if workloads.digest(plan.multiply_operand(a, fixed)) != e["c"]]
    assert not bad, "resident products differ from the oracle: %s" % bad


def test_served_workload_poisson(polymul, workloads):
    """Open-loop Poisson arrivals (100 requests/s, fixed seed) into a server with
    a 200 GB budget: requests are admitted while others run; every result must
    be the oracle's."""
    import time

    g = GOLDEN["served"]["products"]
    products = workloads.served_products()
    _announce("served workload", len(products))
    for (n, limbs, q, a, b), e in zip(products, g):
        _check_inputs(workloads, e, q, a, b, "served request N=%d L=%d" % (n, limbs))
    arrivals = np.cumsum(np.random.default_rng(0x5EED0007).exponential(1.0 / 100.0, size=len(products)))
    with polymul.Server(200 * GiB, lanes=4) as server:
        start = time.perf_counter()
        tickets = []
        for at, prod in zip(arrivals, products):
            delay = start + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(server.submit(*prod))
        results = [server.wait(t) for t in tickets]
    bad = [i for i, ((c, _), e) in enumerate(zip(results, g)) if workloads.digest(c) != e["c"]]
    assert not bad, "served results differ from the oracle: %s" % bad
    assert all(times["total_ms"] >= times["service_ms"] > 0 for _, times in results)


def test_reference_stress_seeds(polymul, workloads, oracle):
    cases = GOLDEN["stress"]
    _announce("stress seeds", len(cases))
    plans = {}
    try:
        for e in cases:
            n, width, seed = e["n"], e["width"], e["seed"]
            limbs = (width + 31) // 32
            key = (n, width)
            if key not in plans:
                q = oracle.modulus(n, width)
                plans[key] = (q, polymul.Plan(n, width, limbs, oracle.int_to_words(q, limbs)))
            q, plan = plans[key]
            a_list, b_list = oracle.random_pair(n, q, seed)
            a, b = oracle.to_limbs(a_list, limbs), oracle.to_limbs(b_list, limbs)
            _check_inputs(workloads, e, oracle.int_to_words(q, limbs), a, b, "stress %s seed %d" % (key, seed))
            assert workloads.digest(plan.multiply(a, b)) == e["c"], "stress %s seed %d" % (key, seed)
    finally:
        for _, plan in plans.values():
            plan.close()
