"""Batched and served forms: many products, each with its own point and q.
Every result must be bit-identical to the same product run alone through the
plain path (and, for small sizes, to the oracle), whatever else is in flight,
under admission by device memory, and with failing requests in the mix."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GiB = 1 << 30


def _product(workloads, n, limbs, seed):
    rng = np.random.default_rng(seed)
    q = workloads.random_modulus(limbs, rng)
    a = workloads.uniform_below(q, n, rng)
    b = workloads.uniform_below(q, n, rng)
    return n, limbs, q, a, b


def _alone(polymul, n, limbs, q, a, b):
    with polymul.Plan(n, 32 * limbs, limbs, q) as plan:
        return plan.multiply(a, b)


MIX = [(64, 1), (1024, 8), (4096, 3), (2, 2), (16384, 40), (1024, 100), (4096, 16), (256, 5), (8192, 64),
       (1, 4), (32768, 12), (512, 33)]


def test_batch_equals_products_alone(polymul, workloads, oracle):
    print("[gpu-case] batch mix", flush=True)
    products = [_product(workloads, n, limbs, 9000 + i) for i, (n, limbs) in enumerate(MIX)]
    got = polymul.multiply_batch(products, 4 * GiB, lanes=3)
    for (n, limbs, q, a, b), c in zip(products, got):
        assert np.array_equal(c, _alone(polymul, n, limbs, q, a, b)), (n, limbs)
        if n * limbs <= 8192:
            assert np.array_equal(c, oracle.kronecker_gmp(a, b, q)), (n, limbs)


def test_server_results_in_a_mix_equal_results_alone(polymul, workloads):
    print("[gpu-case] server mix", flush=True)
    products = [_product(workloads, n, limbs, 9100 + i) for i, (n, limbs) in enumerate(MIX)]
    expect = [_alone(polymul, *p) for p in products]
    with polymul.Server(4 * GiB, lanes=4) as server:
        # submit in two waves so the second arrives while the first is running
        tickets = [server.submit(*p) for p in products[:6]]
        first = [server.wait(t) for t in tickets[:2]]
        tickets += [server.submit(*p) for p in products[6:]]
        results = first + [server.wait(t) for t in tickets[2:]]
    for (c, times), e, (n, limbs, *_rest) in zip(results, expect, products):
        assert np.array_equal(c, e), (n, limbs)
        assert times["total_ms"] >= times["service_ms"] >= 0.0


def test_admission_by_memory_keeps_results_exact(polymul, workloads):
    """A budget that holds about one N=65536, L=55 request at a time: the others
    wait in the queue for memory, and every result stays exact."""
    print("[gpu-case] server admission", flush=True)
    products = [_product(workloads, 65536, 55, 9200 + i) for i in range(4)]
    with polymul.Plan(65536, 32 * 55, 55, products[0][2]) as plan:
        footprint = plan.info()["device_bytes"]
    expect = [_alone(polymul, *p) for p in products]
    with polymul.Server(int(footprint * 1.5), lanes=3) as server:
        tickets = [server.submit(*p) for p in products]
        results = [server.wait(t) for t in tickets]
    for (c, _), e in zip(results, expect):
        assert np.array_equal(c, e)
    # with room for one request at a time, the later ones queued for memory
    assert max(times["queue_ms"] for _, times in results) > 0.5 * min(times["service_ms"] for _, times in results)


def test_request_larger_than_budget_fails_alone(polymul, workloads):
    print("[gpu-case] server too large + invalid", flush=True)
    big = _product(workloads, 65536, 55, 9300)
    small = _product(workloads, 1024, 8, 9301)
    with polymul.Server(64 << 20, lanes=2) as server:
        t_big = server.submit(*big)
        t_small = server.submit(*small)
        with pytest.raises(polymul.PolymulError) as err:
            server.wait(t_big)
        assert err.value.status == polymul.PM_ERR_OUT_OF_MEMORY
        c, _ = server.wait(t_small)
    assert np.array_equal(c, _alone(polymul, *small))


def test_invalid_request_fails_alone(polymul, workloads):
    n, limbs, q, a, b = _product(workloads, 1024, 8, 9400)
    with polymul.Server(1 * GiB, lanes=2) as server:
        bad = server.submit(1000, limbs, q, a[:1000], b[:1000])  # N not a power of two
        good = server.submit(n, limbs, q, a, b)
        with pytest.raises(polymul.PolymulError) as err:
            server.wait(bad)
        assert err.value.status == polymul.PM_ERR_INVALID_ARGUMENT
        c, _ = server.wait(good)
    assert np.array_equal(c, _alone(polymul, n, limbs, q, a, b))


def test_a_ticket_is_waited_once(polymul, workloads):
    """Two threads waiting on one ticket: one gets the result, the other
    PM_ERR_INVALID_ARGUMENT (whichever order they arrive in), and a later wait
    on the same ticket is refused as well."""
    import threading

    print("[gpu-case] server double wait", flush=True)
    n, limbs, q, a, b = _product(workloads, 65536, 55, 9500)
    lib = polymul.load_library()
    with polymul.Server(1 * GiB, lanes=1) as server:
        ticket = server.submit(n, limbs, q, a, b)
        statuses = []
        threads = [threading.Thread(target=lambda: statuses.append(lib.pm_server_wait(server.handle, ticket, None)))
                   for _ in range(2)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        assert sorted(statuses) == [polymul.PM_OK, polymul.PM_ERR_INVALID_ARGUMENT]
        _, c, _ = server._pending.pop(ticket)  # waited through the C ABI above
        with pytest.raises(polymul.PolymulError) as err:
            server.wait(ticket)
        assert err.value.status == polymul.PM_ERR_INVALID_ARGUMENT
    assert np.array_equal(c, _alone(polymul, n, limbs, q, a, b))
