"""Resident-operand products (b transformed once, kept on the device) must be
bit-identical to the plain product and to the oracle, for several a against
one b, through the host and the device entry points."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,count", [("small", 4), ("mid", 2)])
def test_resident_operand_matches_plain_product(polymul, workloads, oracle, name, count):
    print("[gpu-case] resident", name, flush=True)
    n, width, limbs, q, a, b = workloads.seeded_case(name)
    with polymul.Plan(n, width, limbs, q) as plan:
        with plan.operand(b) as fixed:
            for seed in range(count):
                x, _ = workloads.operands_for(n, q, 1000 + seed)
                got = plan.multiply_operand(x, fixed)
                assert np.array_equal(got, plan.multiply(x, b))
                if name == "small":
                    assert np.array_equal(got, oracle.kronecker_gmp(x, b, q))


def test_resident_operand_device_entry(polymul, workloads, device_memory):
    print("[gpu-case] resident device", flush=True)
    n, width, limbs, q, a, b = workloads.seeded_case("small")
    with polymul.Plan(n, width, limbs, q) as plan:
        expect = plan.multiply(a, b)
        with plan.operand(b) as fixed:
            da = device_memory.upload(a)
            dc = device_memory.alloc(a.nbytes)
            try:
                plan.multiply_operand_device(da, fixed, dc)
                got = device_memory.download(dc, a.shape, np.uint32)
            finally:
                device_memory.free(da)
                device_memory.free(dc)
            assert np.array_equal(got, expect)


def test_output_array_reused(polymul, workloads):
    """multiply and multiply_operand write c into a caller's array when given
    one: the same array twice in a row (filled with junk in between) gives the
    freshly allocated result, and a wrong shape, dtype or a read-only array is
    refused before anything runs."""
    print("[gpu-case] out= arrays", flush=True)
    n, width, limbs, q, a, b = workloads.seeded_case("small")
    with polymul.Plan(n, width, limbs, q) as plan:
        expect = plan.multiply(a, b)
        out = np.full((n, limbs), 0xDEADBEEF, dtype=np.uint32)
        for _ in range(2):
            assert plan.multiply(a, b, out=out) is out
            assert np.array_equal(out, expect)
            out.fill(0x5A5A5A5A)
        with plan.operand(b) as fixed:
            assert plan.multiply_operand(a, fixed, out=out) is out
            assert np.array_equal(out, expect)
        bad = [np.empty((n, limbs + 1), dtype=np.uint32), np.empty((n, limbs), dtype=np.int64),
               np.empty((limbs, n), dtype=np.uint32).T]
        ro = np.empty((n, limbs), dtype=np.uint32)
        ro.flags.writeable = False
        for x in bad + [ro]:
            with pytest.raises(ValueError):
                plan.multiply(a, b, out=x)


def test_operand_of_another_plan_is_rejected(polymul, workloads):
    n, width, limbs, q, a, b = workloads.seeded_case("small")
    with polymul.Plan(n, width, limbs, q) as p1, polymul.Plan(n, width, limbs, q) as p2:
        with p1.operand(b) as fixed:
            with pytest.raises(polymul.PolymulError) as err:
                p2.multiply_operand(a, fixed)
            assert err.value.status == 1


@pytest.mark.parametrize("name", ["small", "mid", "large"])
def test_device_entry_matches_host_entry(polymul, workloads, oracle, device_memory, name):
    """pm_multiply_device (operands already in HBM) gives the host entry's c,
    up to the benchmark's own shape (N = 65536, L = 1747)."""
    print("[gpu-case] device entry", name, flush=True)
    n, width, limbs, q, a, b = workloads.seeded_case(name)
    with polymul.Plan(n, width, limbs, q) as plan:
        expect = plan.multiply(a, b)
        da, db = device_memory.upload(a), device_memory.upload(b)
        dc = device_memory.alloc(a.nbytes)
        try:
            for _ in range(2):  # the second call reuses every buffer of the first
                plan.multiply_device(da, db, dc)
                got = device_memory.download(dc, a.shape, np.uint32)
                assert np.array_equal(got, expect)
        finally:
            for ptr in (da, db, dc):
                device_memory.free(ptr)
        if name == "small":
            assert np.array_equal(expect, oracle.kronecker_gmp(a, b, q))


def test_split_entry_matches_multiply(polymul, workloads):
    """pm_multiply_begin / pm_multiply_finish give pm_multiply's c; a second
    begin before finish is refused; finish(NULL) discards and frees the plan."""
    print("[gpu-case] begin/finish", flush=True)
    n, width, limbs, q, a, b = workloads.seeded_case("small")
    with polymul.Plan(n, width, limbs, q) as plan:
        expect = plan.multiply(a, b)
        plan.multiply_begin(a, b)
        with pytest.raises(polymul.PolymulError) as err:
            plan.multiply_begin(a, b)
        assert err.value.status == polymul.PM_ERR_INVALID_ARGUMENT
        assert np.array_equal(plan.multiply_finish(), expect)
        plan.multiply_begin(b, a)
        polymul.load_library().pm_multiply_finish(plan.handle, None)  # discard
        assert np.array_equal(plan.multiply(a, b), expect)


def test_begun_product_blocks_the_plans_other_entries(polymul, workloads, device_memory):
    """Between begin and finish the device entry, operand creation and the
    stage tap are refused (they would reuse the buffers the begun product's
    queued kernels still use); the begun product's result is unaffected."""
    print("[gpu-case] begun product blocks other entries", flush=True)
    n, width, limbs, q, a, b = workloads.seeded_case("small")
    with polymul.Plan(n, width, limbs, q) as plan:
        expect = plan.multiply(a, b)
        da, db = device_memory.upload(b), device_memory.upload(a)
        dc = device_memory.alloc(a.nbytes)
        try:
            plan.multiply_begin(a, b)
            for call in (lambda: plan.multiply_device(da, db, dc), lambda: plan.operand(b),
                         lambda: plan.debug_stage(a, b, polymul.STAGE_OUTPUT)):
                with pytest.raises(polymul.PolymulError) as err:
                    call()
                assert err.value.status == polymul.PM_ERR_INVALID_ARGUMENT
            assert np.array_equal(plan.multiply_finish(), expect)
            plan.multiply_device(da, db, dc)  # accepted again once finished
            assert np.array_equal(device_memory.download(dc, a.shape, np.uint32), plan.multiply(b, a))
        finally:
            for ptr in (da, db, dc):
                device_memory.free(ptr)


def test_grouped_batch_matches_device_entry(polymul, workloads, oracle, device_memory):
    """pm_multiply_device_batch over a batch of mixed shapes -- the reference
    point's N = 65536 with a 1747-bit q, B-mid (N = 32768, split transform),
    q wider than 62 limbs (wave reconstruction), N = 1024 and 2048 (no column
    stage), and two products that run on their own inside the batch (q = 1,
    N = 8) -- gives every product's pm_multiply_device result; a plan named
    twice, a null operand and plans of an empty batch are handled."""
    print("[gpu-case] grouped batch, mixed shapes", flush=True)
    rng = np.random.default_rng(4242)
    cases = []
    n, width, limbs, q, a, b = workloads.seeded_case("mid")
    cases.append((n, limbs, q, a, b))
    for n, limbs in ((65536, 55), (4096, 80), (1024, 8), (2048, 20), (8, 3)):
        q = workloads.random_modulus(limbs, rng)
        cases.append((n, limbs, q, workloads.uniform_below(q, n, rng), workloads.uniform_below(q, n, rng)))
    one = np.zeros(2, dtype=np.uint32)
    one[0] = 1
    cases.append((64, 2, one, np.zeros((64, 2), dtype=np.uint32), np.zeros((64, 2), dtype=np.uint32)))
    plans = [polymul.Plan(n, 32 * limbs, limbs, q) for n, limbs, q, _a, _b in cases]
    bufs = [(device_memory.upload(a), device_memory.upload(b), device_memory.alloc(a.nbytes))
            for _n, _l, _q, a, b in cases]
    try:
        expect = []
        for plan, (pa, pb, pc), (n, limbs, *_r) in zip(plans, bufs, cases):
            plan.multiply_device(pa, pb, pc)
            expect.append(device_memory.download(pc, (n, limbs), np.uint32))
        n, limbs, q, a, b = cases[4]  # N = 1024: against the oracle too
        assert np.array_equal(expect[4], oracle.kronecker_gmp(a, b, q))
        for attempt in range(2):
            for (n, limbs, *_r), (_pa, _pb, pc) in zip(cases, bufs):
                device_memory.write(pc, np.full((n, limbs), 0xDEADBEEF, dtype=np.uint32))
            polymul.multiply_device_batch(plans, [x[0] for x in bufs], [x[1] for x in bufs], [x[2] for x in bufs])
            for i, ((n, limbs, *_r), (_pa, _pb, pc)) in enumerate(zip(cases, bufs)):
                got = device_memory.download(pc, (n, limbs), np.uint32)
                assert np.array_equal(got, expect[i]), "product %d (N=%d, L=%d), pass %d" % (i, n, limbs, attempt)
        with pytest.raises(polymul.PolymulError) as err:
            polymul.multiply_device_batch(plans[:2] + plans[:1], [bufs[0][0], bufs[1][0], bufs[0][0]],
                                          [bufs[0][1], bufs[1][1], bufs[0][1]], [bufs[0][2], bufs[1][2], bufs[0][2]])
        assert err.value.status == 1
        with pytest.raises(polymul.PolymulError) as err:
            polymul.multiply_device_batch(plans[:1], [0], [bufs[0][1]], [bufs[0][2]])
        assert err.value.status == 1
        polymul.multiply_device_batch([], [], [], [])
    finally:
        for buffers in bufs:
            for ptr in buffers:
                device_memory.free(ptr)
        for plan in plans:
            plan.close()


def test_entry_points_share_a_plan(polymul, workloads, device_memory):
    """One B-mid plan (N = 32768, L = 868: column-strip chunks on the host
    path) used through the host-vector entry, the device entry, the grouped
    batch entry (beside a second plan) and the host entry again, in turn:
    every result equals the first, so no entry leaves state that another
    trips over."""
    print("[gpu-case] entry points share a plan", flush=True)
    n, width, limbs, q, a, b = workloads.seeded_case("mid")
    rng = np.random.default_rng(31)
    q2 = workloads.random_modulus(40, rng)
    a2, b2 = workloads.uniform_below(q2, 4096, rng), workloads.uniform_below(q2, 4096, rng)
    with polymul.Plan(n, width, limbs, q) as plan, polymul.Plan(4096, 32 * 40, 40, q2) as other:
        expect = plan.multiply(a, b)
        expect2 = other.multiply(a2, b2)
        bufs = [(device_memory.upload(x), device_memory.upload(y), device_memory.alloc(x.nbytes))
                for x, y in ((a, b), (a2, b2))]
        try:
            for _ in range(2):
                plan.multiply_device(*bufs[0])
                assert np.array_equal(device_memory.download(bufs[0][2], a.shape, np.uint32), expect)
                device_memory.write(bufs[0][2], np.zeros_like(a))
                polymul.multiply_device_batch([other, plan], [bufs[1][0], bufs[0][0]], [bufs[1][1], bufs[0][1]],
                                              [bufs[1][2], bufs[0][2]])
                assert np.array_equal(device_memory.download(bufs[0][2], a.shape, np.uint32), expect)
                assert np.array_equal(device_memory.download(bufs[1][2], a2.shape, np.uint32), expect2)
                assert np.array_equal(plan.multiply(a, b), expect)
        finally:
            for buffers in bufs:
                for ptr in buffers:
                    device_memory.free(ptr)
