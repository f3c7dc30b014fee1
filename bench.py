"""Benchmark of the gfx950 exact negacyclic multiplier.

Workload (one "step" = one full product, device-resident operands):
  N = 65536, L = 1747 limbs: q a seeded random odd 55,904-bit modulus, a and b
  uniform in [0, q) (458 MB each), c = a*b mod (X^N + 1, q), bit-exact.

Prints ONE JSON line with ms/product (lower is better), the live HIP-event
timing of the dominant kernel (the conversion GEMM on INT8 MFMA) as a
roofline fraction, the transfer-inclusive time through the host-vector entry
point, the reference's own competition point (N=65536, 1747-bit NTT-prime q,
55 limbs) and a bounded CPU GMP baseline run on this host.
"""
import argparse
import importlib.util
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
PKG = ROOT / "fherma-polynomial-cuda_amd"
METRIC = "ms/product at N=65536,W=1747 (bit-exact); batched products/s vs plain; p99"
INT8_PEAK_TOPS = 5000.0  # MI355X dense INT8 MFMA, TOP/s
HBM_PEAK_GBS = 8000.0


def _load(name):
    spec = importlib.util.spec_from_file_location("pm_" + name, PKG / (name + ".py"))
    module = importlib.util.module_from_spec(spec)
    sys.modules["pm_" + name] = module
    spec.loader.exec_module(module)
    return module


def conversion_ops(n, limbs, r):
    """Algorithmic INT8 ops of one conversion launch: both operands, 2 ops per
    byte MAC, unpadded K = 4L bytes, 4 digit rows per prime."""
    return 2 * 2 * n * (4 * limbs) * (4 * r)


def crt_ops(n, limbs, r):
    """Algorithmic INT8 ops of one CRT launch: K = 4r residue bytes, 4(L+2) output bytes."""
    return 2 * n * (4 * r) * (4 * (limbs + 2))


def time_device(plan, torch, a_np, b_np, steps, warmup):
    n, limbs = a_np.shape
    dev = torch.device("cuda:0")
    a = torch.from_numpy(a_np.view(np.int32)).to(dev)
    b = torch.from_numpy(b_np.view(np.int32)).to(dev)
    c = torch.empty((n, limbs), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    plan.set_profiling(True)
    for _ in range(warmup):
        plan.multiply_device(a.data_ptr(), b.data_ptr(), c.data_ptr())
    torch.cuda.synchronize()
    stage_sum = None
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.multiply_device(a.data_ptr(), b.data_ptr(), c.data_ptr())
        times = plan.stage_times()
        stage_sum = {k: v + (stage_sum or {}).get(k, 0.0) for k, v in times.items()}
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    stage_avg = {k: v / steps for k, v in stage_sum.items()}
    out = c.cpu().numpy().view(np.uint32)
    del a, b, c
    return elapsed * 1000.0 / steps, stage_avg, out


def time_host(plan, a, b, runs):
    """Host vectors in, a newly allocated host vector out (pm_multiply, transfers
    and output allocation included, as in the harness's timed fherma_run): one
    untimed call, then `runs` timed calls.  Returns (per-call ms, last c)."""
    c = plan.multiply(a, b)
    times = []
    for _ in range(runs):
        del c
        t0 = time.perf_counter()
        c = plan.multiply(a, b)
        times.append((time.perf_counter() - t0) * 1000.0)
    return times, c


GiB = 1 << 30


def plain_products(polymul, products):
    """One at a time, as a harness would run fherma_init + fherma_run per point:
    a plan for the product's own q, host vectors in and out."""
    outs, times = [], []
    for n, limbs, q, a, b in products:
        t0 = time.perf_counter()
        with polymul.Plan(n, 32 * limbs, limbs, q) as plan:
            outs.append(plan.multiply(a, b))
        times.append((time.perf_counter() - t0) * 1000.0)
    return outs, times


def batched_leg(polymul, workloads, count, lanes, passes=3):
    """64 products: N in {2^12..2^16}, L log-uniform in [16, 1747] limbs, own q
    each; device footprint capped at 60 GB; plain path vs pm_multiply_batch.
    Each form first runs once untimed (pinned staging filled for the first
    time, first launches at each shape: both forms pay these once per process),
    then `passes` timed passes over the same 64 products; products/s is from
    the median pass.  The untimed first passes are reported as *_cold."""
    products = workloads.batch_products(count)

    def timed(fn):
        t0 = time.perf_counter()
        out = fn()
        return time.perf_counter() - t0, out

    plain_cold, (plain, _) = timed(lambda: plain_products(polymul, products))
    plain_s = float(np.median([timed(lambda: plain_products(polymul, products))[0] for _ in range(passes)]))
    batched_cold, batched = timed(lambda: polymul.multiply_batch(products, 60 * GiB, lanes=lanes))
    runs = [timed(lambda: polymul.multiply_batch(products, 60 * GiB, lanes=lanes)) for _ in range(passes)]
    batched_s = float(np.median([s for s, _ in runs]))
    exact = all(np.array_equal(x, y) for x, y in zip(plain, batched)) and all(
        np.array_equal(x, y) for x, y in zip(plain, runs[-1][1]))
    words = sum(n * limbs for n, limbs, *_ in products)
    return {"products": count, "lanes": lanes, "device_budget_gib": 60, "timed_passes": passes,
            "plain_products_per_s": round(count / plain_s, 2), "batched_products_per_s": round(count / batched_s, 2),
            "batched_vs_plain": round(plain_s / batched_s, 2), "exact_vs_plain": exact,
            "plain_cold_products_per_s": round(count / plain_cold, 2),
            "batched_cold_products_per_s": round(count / batched_cold, 2),
            "host_link_gb": round(12 * words / 1e9, 2)}, products


def device_batch_leg(polymul, torch, products, lanes, passes=3, golden_key="batch"):
    """A batched workload with every operand already in HBM, apart from the
    host link and host memory: all 64 plans are built first (device tables)
    and the operands uploaded.  Plain runs pm_multiply_device one product after
    another; batched runs `lanes` host threads, each calling pm_multiply_device
    on its next product (largest first, by N*L^2), so the products' kernel
    sequences overlap on the device through the plans' own streams; grouped
    is one pm_multiply_device_batch call, each stage one launch for all
    products.  Every
    output is hashed and compared with the oracle's digest of the workload
    (tests/golden/workload_digests.json, entry `golden_key`)."""
    import hashlib
    import threading

    golden = json.loads((ROOT / "tests" / "golden" / "workload_digests.json").read_text())[golden_key]["products"]
    dev = torch.device("cuda:0")
    plans, bufs = [], []
    try:
        for n, limbs, q, a, b in products:
            plans.append(polymul.Plan(n, 32 * limbs, limbs, q))
            ta = torch.from_numpy(a.view(np.int32)).to(dev)
            tb = torch.from_numpy(b.view(np.int32)).to(dev)
            bufs.append((ta, tb, torch.empty_like(ta)))
        torch.cuda.synchronize()
        order = sorted(range(len(products)), key=lambda i: -products[i][0] * products[i][1] ** 2)

        def one(i):
            ta, tb, tc = bufs[i]
            plans[i].multiply_device(ta.data_ptr(), tb.data_ptr(), tc.data_ptr())

        def plain():
            for i in order:
                one(i)

        def batched():
            pending = iter(order)
            lock = threading.Lock()

            def worker():
                while True:
                    with lock:
                        i = next(pending, None)
                    if i is None:
                        return
                    one(i)

            threads = [threading.Thread(target=worker) for _ in range(lanes)]
            for t in threads:
                t.start()
            for t in threads:
                t.join()

        def timed(fn):
            fn()  # untimed first pass
            torch.cuda.synchronize()
            runs = []
            for _ in range(passes):
                t0 = time.perf_counter()
                fn()
                torch.cuda.synchronize()
                runs.append(time.perf_counter() - t0)
            return float(np.median(runs))

        def digests():
            out = []
            for i, (n, limbs, *_rest) in enumerate(products):
                c = bufs[i][2].cpu().numpy().view(np.uint32).reshape(n, limbs)
                out.append(hashlib.sha256(np.ascontiguousarray(c, dtype="<u4").tobytes()).hexdigest())
            return out

        pointers = [[t.data_ptr() for t in column] for column in zip(*bufs)]

        def grouped():
            polymul.multiply_device_batch(plans, *pointers)

        plain_s = timed(plain)
        plain_ok = digests() == [g["c"] for g in golden]
        for _, _, tc in bufs:
            tc.zero_()
        batched_s = timed(batched)
        batched_ok = digests() == [g["c"] for g in golden]
        for _, _, tc in bufs:
            tc.zero_()
        grouped_s = timed(grouped)
        grouped_ok = digests() == [g["c"] for g in golden]
        count = len(products)
        return {"products": count, "lanes": lanes, "timed_passes": passes,
                "plain_products_per_s": round(count / plain_s, 2),
                "batched_products_per_s": round(count / batched_s, 2),
                "batched_vs_plain": round(plain_s / batched_s, 2),
                "grouped_products_per_s": round(count / grouped_s, 2),
                "grouped_vs_plain": round(plain_s / grouped_s, 2),
                "plain_ms": round(plain_s * 1000.0, 1), "batched_ms": round(batched_s * 1000.0, 1),
                "grouped_ms": round(grouped_s * 1000.0, 2),
                "exact_vs_oracle_digests": bool(plain_ok and batched_ok and grouped_ok)}
    finally:
        for plan in plans:
            plan.close()
        del bufs


def resident_leg(polymul, workloads, count):
    """64 random a times one fixed b: plain products vs b's residues kept on the
    device (b's transfer, conversion and forward transform skipped).  Timed
    twice: each c into a fresh array, as a one-off caller would, and into one
    array reused, which takes the fresh pages' faults (about as long as a
    product here) out of both sides."""
    n, limbs, q, b, a_list = workloads.resident_case(count)
    with polymul.Plan(n, 32 * limbs, limbs, q) as plan:
        plan.multiply(a_list[0], b)
        t0 = time.perf_counter()
        plain = [plan.multiply(a, b) for a in a_list]
        plain_s = time.perf_counter() - t0
        out = np.empty((n, limbs), dtype=np.uint32)
        t0 = time.perf_counter()
        for a in a_list:
            plan.multiply(a, b, out=out)
        plain_reused_s = time.perf_counter() - t0
        with plan.operand(b) as fixed:
            t0 = time.perf_counter()
            resident = [plan.multiply_operand(a, fixed) for a in a_list]
            resident_s = time.perf_counter() - t0
            t0 = time.perf_counter()
            for a in a_list:
                plan.multiply_operand(a, fixed, out=out)
            resident_reused_s = time.perf_counter() - t0
            last_reused = out.copy()
    exact = all(np.array_equal(x, y) for x, y in zip(plain, resident)) and np.array_equal(last_reused, plain[-1])
    return {"products": count, "n": n, "limbs": limbs, "plain_products_per_s": round(count / plain_s, 2),
            "resident_products_per_s": round(count / resident_s, 2),
            "resident_vs_plain": round(plain_s / resident_s, 2),
            "reused_output": {"plain_products_per_s": round(count / plain_reused_s, 2),
                              "resident_products_per_s": round(count / resident_reused_s, 2),
                              "resident_vs_plain": round(plain_reused_s / resident_reused_s, 2)},
            "exact_vs_plain": exact}


def served_leg(polymul, workloads, count, lanes, rates):
    """Open-loop Poisson arrivals at each rate; N = 2^U, U uniform in 12..16, L
    log-uniform in [32, 1747] limbs, own q each; 200 GB admission budget.
    Latency = submission to result (server clock); every result must equal the
    same request run alone through the plain path."""
    products = workloads.served_products(count)
    alone, alone_ms = plain_products(polymul, products)
    sweeps = []
    for rate in rates:
        rng = np.random.default_rng(0x5EED0007)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=count))
        with polymul.Server(200 * GiB, lanes=lanes) as server:
            start = time.perf_counter()
            tickets = []
            for at, prod in zip(arrivals, products):
                delay = start + at - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                tickets.append((time.perf_counter(), server.submit(*prod)))
            results = [(sub, server.wait(t)) for sub, t in tickets]
        finish = max(sub + times["total_ms"] / 1000.0 for sub, (_, times) in results)
        lat = np.array([times["total_ms"] for _, (_, times) in results])
        exact = all(np.array_equal(c, e) for (_, (c, _)), e in zip(results, alone))
        sweeps.append({"offered_per_s": round(rate, 2), "products_per_s": round(count / (finish - start), 2),
                       "p50_ms": round(float(np.percentile(lat, 50)), 2),
                       "p99_ms": round(float(np.percentile(lat, 99)), 2),
                       "queue_p99_ms": round(float(np.percentile([t["queue_ms"] for _, (_, t) in results], 99)), 2),
                       "exact_vs_alone": exact})
    return {"requests_per_rate": count, "lanes": lanes, "device_budget_gib": 200,
            "plain_one_at_a_time_products_per_s": round(count / (sum(alone_ms) / 1000.0), 2),
            "plain_p50_ms": round(float(np.percentile(alone_ms, 50)), 2),
            "plain_p99_ms": round(float(np.percentile(alone_ms, 99)), 2), "sweep": sweeps}


def harness_leg(polymul, workloads, n, width, limbs, q, cases):
    """End to end through the drop-in: the case runner (fherma-polynomial-cuda_amd/
    solution, a child process) calls fherma_init once and times each fherma_run
    with steady_clock, as the reference's harness does (src/main.cpp:181-186).
    Case 0 is a warm-up; two answers are compared with the plain product."""
    import shutil
    import subprocess
    import tempfile

    root = Path(tempfile.mkdtemp(prefix="pm_point_"))
    try:
        operands = [workloads.operands_for(n, q, 7000 + i) for i in range(cases)]
        workloads.write_point_dir(root, n, width, limbs, q, operands)
        proc = subprocess.run([str(PKG / "solution"), str(root)], capture_output=True, text=True, timeout=300)
        if proc.returncode != 0:
            raise RuntimeError("case runner failed: %s" % proc.stderr[-500:])
        res = json.loads((root / "out" / "results.json").read_text())
        ok = [c for c in res["cases"] if c["status"] == "ok"]
        ms = np.array([c["seconds"] * 1000.0 for c in ok[1:]])
        with polymul.Plan(n, width, limbs, q) as plan:
            exact = all(np.array_equal(np.fromfile(root / "out" / ("%06d" % i) / "c.bin", dtype="<u4").reshape(n, limbs),
                                       plan.multiply(*operands[i])) for i in (0, cases - 1))
        return {"cases_timed": int(ms.size), "median_ms": round(float(np.median(ms)), 3),
                "p99_ms": round(float(np.percentile(ms, 99)), 3), "min_ms": round(float(ms.min()), 3),
                "init_s": round(res["init_s"], 3), "all_ok": len(ok) == cases, "exact_vs_plain": exact}
    finally:
        shutil.rmtree(root, ignore_errors=True)


def host_copy_rate(threads=(1, 4, 8, 16), gib=1):
    """Host DRAM copy rate of this machine, the resource the host-vector path
    spends besides the link: numpy copies between two `gib` GiB arrays, both
    touched beforehand, by 1..16 threads over disjoint slices (numpy drops the
    GIL inside a copy).  GB/s counts bytes read plus bytes written; best of 3."""
    import threading

    src = np.ones(gib << 28, dtype=np.uint32)
    dst = np.zeros_like(src)
    rates = {}
    for t in threads:
        bounds = np.linspace(0, src.size, t + 1).astype(np.int64)

        def work(i):
            np.copyto(dst[bounds[i]:bounds[i + 1]], src[bounds[i]:bounds[i + 1]])

        best = None
        for _ in range(3):
            workers = [threading.Thread(target=work, args=(i,)) for i in range(t)]
            t0 = time.perf_counter()
            for w in workers:
                w.start()
            for w in workers:
                w.join()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        rates["threads_%d_gbs" % t] = round(2 * src.nbytes / best / 1e9, 1)
    del src, dst
    return rates


def cpu_baseline(q, a, b, plan_cls, sample_n):
    """The oracle's GMP Kronecker product (one core) on a bounded sample: the
    first sample_n coefficients of each operand as an N=sample_n product with
    the same q; the GPU computes the same sample for an exact comparison."""
    sys.path.insert(0, str(ROOT))
    from oracle import reference

    sa = np.ascontiguousarray(a[:sample_n])
    sb = np.ascontiguousarray(b[:sample_n])
    t0 = time.perf_counter()
    expect = reference.kronecker_gmp(sa, sb, q)
    cpu_s = time.perf_counter() - t0
    limbs = q.size
    with plan_cls(sample_n, 32 * limbs, limbs, q) as plan:
        got = plan.multiply(sa, sb)
    exact = bool(np.array_equal(got, expect))
    # GMP's product of Kronecker integers of B bits costs ~B log B; scale the
    # sample (sample_n coefficients) to the full N by that law.
    bits_sample = sample_n * ((sample_n * (reference.words_to_int(q) - 1) ** 2).bit_length())
    n_full = a.shape[0]
    bits_full = n_full * ((n_full * (reference.words_to_int(q) - 1) ** 2).bit_length())
    scale = (bits_full * math.log2(bits_full)) / (bits_sample * math.log2(bits_sample))
    return cpu_s, scale, exact, reference


def calibrate():
    """Register-only throughput of the instructions the rooflines are built on
    (libpm_calibrate.so): INT8 MFMA 16x16x64 and 32x32x32 with every operand in
    registers, and the 32-bit multiply rate (v_mul_lo_u32 + v_mul_hi_u32) that
    bounds the NTT row stage.  256 CUs x 8 blocks, 8 waves per block."""
    import ctypes

    lib = ctypes.CDLL(str(PKG / "libpm_calibrate.so"))
    ms, ops = ctypes.c_float(), ctypes.c_double()
    out = {}
    for shape in (16, 32):
        rc = lib.pm_calibrate_mfma_i8(shape, 256 * 8, 8, 2048, ctypes.byref(ms), ctypes.byref(ops))
        if rc == 0:
            out["mfma_i8_%dx%d_tops" % (shape, shape)] = round(ops.value / (ms.value * 1e-3) / 1e12, 1)
    if lib.pm_calibrate_mul32(256 * 16, 4096, ctypes.byref(ms), ctypes.byref(ops)) == 0:
        out["mul32_per_s"] = float("%.4g" % (ops.value / (ms.value * 1e-3)))
    return out


def _oracle_worker(name):
    """Child process: one GMP Kronecker product (oracle, 1 core) of the seeded
    workload `name`; prints its monotonic start and end times as JSON."""
    sys.path.insert(0, str(ROOT))
    from oracle import reference

    workloads = _load("workloads")
    n, width, limbs, q, a, b = workloads.seeded_case(name)
    t0 = time.monotonic()
    reference.kronecker_gmp(a, b, q)
    print(json.dumps({"start": t0, "end": time.monotonic()}), flush=True)


def _oracle_processes(name, count):
    """`count` concurrent oracle products of workload `name`, one process (and
    core) each; returns (seconds of each, wall seconds from first start to last end)."""
    import subprocess

    procs = [subprocess.Popen([sys.executable, str(ROOT / "bench.py"), "--oracle-worker", name],
                              stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True) for _ in range(count)]
    spans = []
    for proc in procs:
        out, err = proc.communicate(timeout=900)
        if proc.returncode != 0:
            raise RuntimeError("oracle worker failed: %s" % err[-500:])
        spans.append(json.loads(out.strip().splitlines()[-1]))
    wall = max(x["end"] for x in spans) - min(x["start"] for x in spans)
    return [x["end"] - x["start"] for x in spans], wall


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


PROFILE_ARGS = "--steps 1 --warmup 0 --no-cpu --no-secondary --no-batch"
SQ_COUNTERS = ["SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "SQ_VALU_MFMA_BUSY_CYCLES", "SQ_LDS_BANK_CONFLICT",
               "SQ_LDS_IDX_ACTIVE", "SQ_WAIT_INST_LDS", "SQ_INSTS_LDS", "SQ_INSTS_VALU_MFMA_MOPS_I8"]


def _short_kernel(name):
    return name.replace("pm::", "").replace("(anonymous namespace)::", "").split("(")[0]


def _counters(path):
    """rocprofv3 counter_collection.csv -> {kernel: {counter: [value per dispatch]}}."""
    import csv
    from collections import defaultdict

    values = defaultdict(lambda: defaultdict(list))
    with open(path) as f:
        for row in csv.DictReader(f):
            values[_short_kernel(row["Kernel_Name"])][row["Counter_Name"]].append(float(row["Counter_Value"]))
    return values


def summarize_traffic(fetch_csv, write_csv, dest):
    """HBM bytes per dispatch from two counter runs of this script with
    --calibrate-copy (FETCH_SIZE, then WRITE_SIZE; both in KiB). Each counter is
    scaled by what it reads for the calibration copy, a plain 1 GiB
    device-to-device copy (the largest copyBuffer dispatch): 1 GiB read, 1 GiB
    written."""
    copy_bytes = 1 << 30
    fetch = {k: [x * 1024.0 for x in v["FETCH_SIZE"]] for k, v in _counters(fetch_csv).items() if "FETCH_SIZE" in v}
    write = {k: [x * 1024.0 for x in v["WRITE_SIZE"]] for k, v in _counters(write_csv).items() if "WRITE_SIZE" in v}
    copy = next(k for k in fetch if "copyBuffer" in k)
    fetch_scale = copy_bytes / max(fetch[copy])
    write_scale = copy_bytes / max(write[copy])
    out = {"command": "rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE) --output-format csv -- python3 bench.py "
                      + PROFILE_ARGS + " --calibrate-copy",
           "summary": "python3 bench.py --summarize-traffic <FETCH_SIZE counter_collection.csv> "
                      "<WRITE_SIZE counter_collection.csv>",
           "calibration": {"copy_bytes": copy_bytes, "fetch_scale": fetch_scale, "write_scale": write_scale},
           "kernels": {}}
    for name in sorted(set(fetch) & set(write)):
        if "copyBuffer" in name:
            continue
        f = sum(fetch[name]) / len(fetch[name]) * fetch_scale
        w = sum(write[name]) / len(write[name]) * write_scale
        out["kernels"][name] = {"fetch_bytes": round(f), "write_bytes": round(w), "hbm_bytes": round(f + w)}
    Path(dest).write_text(json.dumps(out, indent=1) + "\n")


def _clocks(path, xcds=8):
    """GPU clock during each dispatch from a `--pmc GRBM_GUI_ACTIVE --kernel-trace`
    run: GRBM_GUI_ACTIVE counts busy cycles summed over the XCDs, so the clock
    is that count / XCDs / the dispatch's duration.  -> {kernel: (GHz, ms)}."""
    import csv
    from collections import defaultdict

    runs = defaultdict(list)
    with open(path) as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] != "GRBM_GUI_ACTIVE":
                continue
            seconds = (int(row["End_Timestamp"]) - int(row["Start_Timestamp"])) * 1e-9
            if seconds > 0:
                runs[_short_kernel(row["Kernel_Name"])].append((float(row["Counter_Value"]) / xcds / seconds, seconds))
    return {k: (sum(g for g, _ in v) / len(v) / 1e9, sum(t for _, t in v) / len(v) * 1e3) for k, v in runs.items()}


VALU_COUNTERS = ["SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_BUSY_CYCLES", "SQ_WAVE_CYCLES"]


def summarize_kernels(stats_csv, pmc_csv, dest, trace_dir, pmc_dir, clock_csv=None, valu_csv=None):
    """Markdown table of a --kernel-trace --stats run of the default product
    (5 timed steps after 1 warm-up), plus the per-dispatch SQ counters of a
    one-step counter run when pmc_csv is given, and the clock each kernel ran at
    when clock_csv (a GRBM_GUI_ACTIVE counter run) is given."""
    import csv

    steps = "--steps 5 --warmup 1 --no-cpu --no-secondary --no-batch"
    lines = ["# Kernel summary: plain product N=65536, L=1747 (55,904-bit random odd q)", "",
             "Command (from the repo root, on one MI355X):", "", "```",
             f"rocprofv3 --kernel-trace --stats --output-format csv -d {trace_dir} -o large -- python3 bench.py {steps}",
             "```", "",
             "6 products (1 warm-up + 5 timed). Source: `large_kernel_stats.csv`, summarised by "
             "`python3 bench.py --summarize-kernels <stats csv> [<counter csv>]`. \"timed avg\" is the mean "
             "of the 5 timed products' launches (from `large_kernel_trace.csv`, the warm-up product's "
             "launch dropped), the figure to set beside the bench line's event time.", "",
             "| kernel | calls | avg ms | timed avg ms | min ms | total ms | % |",
             "|---|---|---|---|---|---|---|"]
    per_call = {}
    trace_csv = Path(stats_csv).parent / "large_kernel_trace.csv"
    if trace_csv.is_file():
        with open(trace_csv) as f:
            for r in sorted(csv.DictReader(f), key=lambda r: int(r["Start_Timestamp"])):
                per_call.setdefault(r["Kernel_Name"], []).append(
                    (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6)
    with open(stats_csv) as f:
        for r in csv.DictReader(f):
            calls = per_call.get(r["Name"], [])
            timed = "%.4f" % (sum(calls[1:]) / len(calls[1:])) if len(calls) == 6 else "-"
            lines.append("| %s | %s | %.4f | %s | %.4f | %.3f | %.2f |" % (
                _short_kernel(r["Name"]), r["Calls"], float(r["AverageNs"]) / 1e6, timed, float(r["MinNs"]) / 1e6,
                float(r["TotalDurationNs"]) / 1e6, float(r["Percentage"])))
    if pmc_csv:
        shown = ["SQ_VALU_MFMA_BUSY_CYCLES", "SQ_BUSY_CYCLES", "SQ_INSTS_LDS", "SQ_LDS_IDX_ACTIVE",
                 "SQ_LDS_BANK_CONFLICT", "SQ_WAIT_INST_LDS"]
        lines += ["", "## SQ counters (one product, per dispatch)", "", "```",
                  f"rocprofv3 --pmc {' '.join(SQ_COUNTERS)} --output-format csv -d {pmc_dir} -o large -- "
                  f"python3 bench.py {PROFILE_ARGS}",
                  "```", "", "| kernel | " + " | ".join(shown) + " | conflict / LDS active |",
                  "|---|" + "---|" * (len(shown) + 1)]
        for k, d in _counters(pmc_csv).items():
            if "copyBuffer" in k or "elementwise" in k:
                continue
            avg = {c: sum(v) / len(v) for c, v in d.items()}
            act = avg.get("SQ_LDS_IDX_ACTIVE", 0)
            ratio = "%.2f" % (avg.get("SQ_LDS_BANK_CONFLICT", 0) / act) if act else "-"
            lines.append("| %s | " % k + " | ".join("%.4g" % avg.get(c, 0) for c in shown) + " | %s |" % ratio)
    if clock_csv:
        clk_dir = str(Path(clock_csv).parent)
        lines += ["", "## Clock under load (one product, per dispatch)", "", "```",
                  f"rocprofv3 --pmc GRBM_GUI_ACTIVE GRBM_COUNT --kernel-trace --output-format csv -d {clk_dir} -o large "
                  f"-- python3 bench.py {PROFILE_ARGS}",
                  "```", "",
                  "GHz = GRBM_GUI_ACTIVE (busy cycles summed over the 8 XCDs) / 8 / dispatch duration. For the "
                  "GEMMs and the register-only MFMA calibration kernels, ops per cycle per SIMD = INT8 ops / (duration x "
                  "clock) / 1024 SIMDs, against 2048 for one v_mfma_i32_16x16x64_i8 every 16 cycles (the 5.0 POP/s "
                  "datasheet peak assumes 2.4 GHz). Counter collection itself slows the GEMMs somewhat.",
                  "", "| kernel | GHz | ms (counter run) | INT8 ops / cycle / SIMD |", "|---|---|---|---|"]
        gemm_ops = getattr(summarize_kernels, "gemm_ops", {})
        for k, (ghz, ms) in sorted(_clocks(clock_csv).items(), key=lambda kv: -kv[1][1]):
            if "copyBuffer" in k or "elementwise" in k:
                continue
            ops = next((v for name, v in gemm_ops.items() if name in k), None)
            per = "%.0f" % (ops / (ms * 1e-3 * ghz * 1e9) / 1024) if ops else "-"
            lines.append("| %s | %.2f | %.3f | %s |" % (k, ghz, ms, per))
    if valu_csv and clock_csv:
        clocks = _clocks(clock_csv)
        lines += ["", "## VALU issue (one product, per dispatch)", "", "```",
                  f"rocprofv3 --pmc {' '.join(VALU_COUNTERS)} --output-format csv -d {Path(valu_csv).parent} -o large "
                  f"-- python3 bench.py {PROFILE_ARGS}",
                  "```", "",
                  "VALU share = SQ_INSTS_VALU x 4 cycles (a wave64 VALU instruction on a 16-lane SIMD) / "
                  "(1024 SIMDs x dispatch cycles at the clock above): the fraction of the SIMDs' VALU issue the "
                  "kernel uses.", "", "| kernel | SQ_INSTS_VALU | SQ_ACTIVE_INST_VALU | VALU share |", "|---|---|---|---|"]
        for k, d in _counters(valu_csv).items():
            if "copyBuffer" in k or "elementwise" in k or k not in clocks:
                continue
            ghz, ms = clocks[k]
            insts = sum(d["SQ_INSTS_VALU"]) / len(d["SQ_INSTS_VALU"])
            active = sum(d["SQ_ACTIVE_INST_VALU"]) / len(d["SQ_ACTIVE_INST_VALU"])
            share = insts * 4 / (1024 * ms * 1e-3 * ghz * 1e9)
            lines.append("| %s | %.4g | %.4g | %.2f |" % (k, insts, active, share))
    Path(dest).write_text("\n".join(lines) + "\n")


def summarize_grouped(trace_csv, dest):
    """Markdown summary of a --kernel-trace run of `bench.py --small-batch-only`:
    the 64-product small mix run one product at a time (plain, then from four
    threads: 4 passes each, the first untimed) and as one
    pm_multiply_device_batch call (4 passes).  The grouped kernels are the ones
    named *_group*; the others belong to the per-product passes."""
    import csv

    rows = list(csv.DictReader(open(trace_csv)))
    disp = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"]), _short_kernel(r["Kernel_Name"])) for r in rows]
    # plan building (table_*, fill_zero) and torch's own kernels are neither form
    disp = [d for d in disp if not any(x in d[2] for x in ("copyBuffer", "elementwise", "fill", "table_"))]
    grouped = sorted(d for d in disp if "_group" in d[2])
    single = [d for d in disp if "_group" not in d[2]]
    passes_grouped, passes_single = 4, 8
    per = len(grouped) // passes_grouped
    spans = []
    for k in range(passes_grouped):
        part = grouped[k * per:(k + 1) * per]
        spans.append((part[-1][1] - part[0][0]) / 1e6)
    busy_g = sum(e - s for s, e, _ in grouped) / 1e6 / passes_grouped
    busy_s = sum(e - s for s, e, _ in single) / 1e6 / passes_single
    lines = ["# Small-product mix: one launch per product and stage against one per stage", "",
             "Command (from the repo root, on one MI355X):", "", "```",
             f"rocprofv3 --kernel-trace --output-format csv -d {trace_csv.parent} -o small -- "
             "python3 bench.py --small-batch-only", "```", "",
             "The 64 products of `workloads.batch_small_products()` (N = 2^10..2^13, L log-uniform in [8, 128], "
             "each with its own q), operands in HBM. `pm_multiply_device` one product at a time and from four "
             "threads (4 passes each, the first untimed), then `pm_multiply_device_batch` (4 passes). Summarised by "
             "`python3 bench.py --summarize-grouped <kernel trace csv>`.", "",
             "| form | kernel dispatches per pass | kernel time per pass (ms, summed) |", "|---|---|---|",
             "| one product at a time (plain and four threads) | %d | %.3f |" % (len(single) // passes_single, busy_s),
             "| pm_multiply_device_batch | %d | %.3f |" % (per, busy_g), "",
             "Grouped passes, first kernel start to last kernel end (ms): " + ", ".join("%.3f" % x for x in spans),
             "", "| grouped kernel | dispatches per pass | avg ms |", "|---|---|---|"]
    by = {}
    for s_, e, name in grouped:
        by.setdefault(name, []).append((e - s_) / 1e6)
    for name, v in sorted(by.items(), key=lambda kv: -sum(kv[1])):
        lines.append("| %s | %d | %.4f |" % (name, len(v) // passes_grouped, sum(v) / len(v)))
    Path(dest).write_text("\n".join(lines) + "\n")


def main():
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=10)
    parser.add_argument("--warmup", type=int, default=2)
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    parser.add_argument("--no-secondary", action="store_true", help="skip the reference-point and e2e legs")
    parser.add_argument("--cpu-sample", type=int, default=8192)
    parser.add_argument("--no-batch", action="store_true", help="skip the batched, resident and served legs")
    parser.add_argument("--batch-count", type=int, default=64)
    parser.add_argument("--serve-count", type=int, default=48)
    parser.add_argument("--lanes", type=int, default=4)
    parser.add_argument("--oracle-worker", metavar="CONFIG", help=argparse.SUPPRESS)
    parser.add_argument("--no-calibrate", action="store_true", help="skip the register-only peak kernels")
    parser.add_argument("--calibrate-copy", action="store_true",
                        help="also run one plain 1 GiB device-to-device copy (counter calibration)")
    parser.add_argument("--summarize-traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                        help="write profiles/traffic_large.json from two counter runs, then exit (no GPU)")
    parser.add_argument("--small-batch-only", action="store_true",
                        help="run only the device-resident small-product mix (plain, threads, grouped) and exit")
    parser.add_argument("--summarize-grouped", metavar="TRACE_CSV",
                        help="write profiles/grouped_small.md from a --kernel-trace of --small-batch-only, then "
                             "exit (no GPU)")
    parser.add_argument("--summarize-kernels", nargs="+", metavar="CSV",
                        help="write profiles/kernels_large.md from a kernel-stats csv and optionally an "
                             "SQ counter csv, a GRBM_GUI_ACTIVE counter csv and a VALU counter csv, then exit "
                             "(no GPU)")
    args = parser.parse_args()
    if args.oracle_worker:
        _oracle_worker(args.oracle_worker)
        return
    if args.summarize_traffic:
        summarize_traffic(*args.summarize_traffic, ROOT / "profiles" / "traffic_large.json")
        return
    if args.summarize_grouped:
        summarize_grouped(Path(args.summarize_grouped), ROOT / "profiles" / "grouped_small.md")
        return
    if args.summarize_kernels:
        stats = Path(args.summarize_kernels[0])
        pmc = Path(args.summarize_kernels[1]) if len(args.summarize_kernels) > 1 else None
        clock = Path(args.summarize_kernels[2]) if len(args.summarize_kernels) > 2 else None
        valu = Path(args.summarize_kernels[3]) if len(args.summarize_kernels) > 3 else None
        # algorithmic ops of the two GEMMs at the default workload (3636 primes)
        # and of the register-only calibration launches (calibrate(): 2048 blocks x 8 waves x 2048
        # iterations x 16 chains of 16x16x64, or x 8 chains of 32x32x32)
        summarize_kernels.gemm_ops = {"conversion_gemm": conversion_ops(65536, 1747, 3636),
                                      "crt_gemm": crt_ops(65536, 1747, 3636),
                                      "mfma16_peak": 2048 * 8 * 2048 * 16 * 2 * 16 * 16 * 64,
                                      "mfma32_peak": 2048 * 8 * 2048 * 8 * 2 * 32 * 32 * 32}
        # the -d directories in the recorded commands are those the csv files came from
        summarize_kernels(stats, pmc, ROOT / "profiles" / "kernels_large.md",
                          str(stats.parent), str(pmc.parent) if pmc else None, clock, valu)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: one product runs on one GPU; the multi-GPU "
              "sharding (by primes, one reduce-scatter of CRT word sums) is a design, not a "
              "measured path. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    # The library first: it binds the process to the HIP runtime it was built
    # against (ROCm's libamdhip64.so.7), which torch then shares, instead of
    # torch's bundled copy.  The two runtimes differ in how they move device ->
    # pinned-host copies (ROCm's uses the SDMA engines; torch's copy runs them as
    # blit kernels on the CUs, beside the CRT GEMM chunks they overlap).
    polymul = _load("polymul")
    polymul.load_library()
    import torch

    if args.calibrate_copy:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE counters give
        # the scale that turns those counters into HBM bytes
        src = torch.empty(1 << 28, dtype=torch.int32, device="cuda:0").fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    workloads = _load("workloads")
    if args.small_batch_only:
        # the small-product mix alone (for a kernel trace of its three forms)
        leg = device_batch_leg(polymul, torch, workloads.batch_small_products(), args.lanes,
                               golden_key="batch_small")
        print(json.dumps({"batched_device_small": leg}))
        return
    t_gen = time.perf_counter()
    n, width, limbs, q, a, b = workloads.seeded_case("large")
    gen_s = time.perf_counter() - t_gen
    t_plan = time.perf_counter()
    plan = polymul.Plan(n, width, limbs, q)
    plan_s = time.perf_counter() - t_plan
    info = plan.info()
    r = info["primes"]

    ms, stages, c = time_device(plan, torch, a, b, args.steps, args.warmup)

    result = {
        "metric": METRIC,
        "value": round(ms, 4),
        "unit": "ms/product",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded uniform operands below a seeded random odd q)",
        "config": {"workload": "plain product, N=65536, L=1747 limbs (55,904-bit random odd q), "
                               "operands resident in HBM",
                   "n": n, "limbs": limbs, "primes": r, "split": info["block"], "transform": info["transform"],
                   "parallelism": "single GPU"},
    }
    conv_ms = stages["conversion_gemm"]
    ops = conversion_ops(n, limbs, r)
    achieved = ops / (conv_ms * 1e-3) / 1e12
    traffic = None
    traffic_file = ROOT / "profiles" / "traffic_large.json"
    if traffic_file.is_file():  # measured HBM bytes per dispatch (rocprofv3 FETCH_SIZE + WRITE_SIZE)
        entry = json.loads(traffic_file.read_text())["kernels"].get("conversion_gemm")
        traffic = entry["hbm_bytes"] if entry else None
    result["roofline"] = {"kernel": "conversion_gemm", "bound": "mfma", "achieved": round(achieved, 1),
                          "peak": INT8_PEAK_TOPS, "unit": "TOP/s", "frac": round(achieved / INT8_PEAK_TOPS, 4),
                          "traffic": traffic, "ops_per_launch": ops, "avg_launch_ms": round(conv_ms, 4)}
    if not args.no_calibrate:
        peaks = calibrate()
        result["measured_peaks"] = peaks
        measured = peaks.get("mfma_i8_16x16_tops")
        if measured:
            result["roofline"]["measured_peak"] = measured
            result["roofline"]["frac_of_measured_peak"] = round(achieved / measured, 4)
    result["stages_ms"] = {k: round(v, 4) for k, v in stages.items()}
    result["hip_runtime"] = sorted(polymul.hip_runtimes())
    # every output word of the timed product against the oracle's GMP product of
    # the same seeded inputs (its SHA-256, computed once and committed)
    golden = json.loads((ROOT / "tests" / "golden" / "workload_digests.json").read_text())["large"]
    result["exact_vs_oracle_digest"] = workloads.digest(c) == golden["c"] and workloads.digest(a) == golden["a"]
    result["crt_gemm_tops"] = round(crt_ops(n, limbs, r) / (stages["crt_gemm"] * 1e-3) / 1e12, 1)
    result["setup_s"] = {"inputs": round(gen_s, 2), "plan": round(plan_s, 2)}

    if not args.no_secondary:
        # transfer-inclusive: host vectors in, host vector out (pm_multiply)
        plan.set_profiling(False)
        e2e, c_host = time_host(plan, a, b, 3)
        result["e2e_host_ms"] = {"median": round(float(np.median(e2e)), 3), "min": round(min(e2e), 3),
                                 "runs": len(e2e)}
        result["host_copy_rate"] = host_copy_rate()
        result["e2e_matches_device"] = bool(np.array_equal(c_host, c))
        del c_host
    plan.close()

    if not args.no_secondary:
        points = json.loads((ROOT / "tests" / "golden" / "points.json").read_text())
        ref = points["65536/1747"]
        ref_q = workloads.int_to_words(int(ref["q"], 16), ref["L"])
        ra, rb = workloads.operands_for(ref["N"], ref_q, 101)
        with polymul.Plan(ref["N"], ref["W"], ref["L"], ref_q) as rplan:
            rms, rstages, _ = time_device(rplan, torch, ra, rb, max(args.steps, 10), args.warmup)
        # end to end as the harness measures it: fherma_run timed by the case runner
        # (a separate process, so this one's HIP runtime does not enter the figure)
        harness = harness_leg(polymul, workloads, ref["N"], ref["W"], ref["L"], ref_q, 21)
        result["reference_point"] = {"N": 65536, "W_bits": 1747, "L": 55, "ms_device": round(rms, 4),
                                     "fherma_run_ms": harness,
                                     "published_ms_e2e_rtx_pro_6000": 1.0712435}
        # the smaller configurations (N=1024 with a 256-bit q, N=32768 with a
        # 27,776-bit q): kernel-only and transfer-inclusive ms per product
        others = {}
        for name, label in (("small", "B-small"), ("mid", "B-mid")):
            sn, sw, sl, sq, sa, sb = workloads.seeded_case(name)
            with polymul.Plan(sn, sw, sl, sq) as splan:
                dms, _, dc = time_device(splan, torch, sa, sb, max(args.steps, 10), args.warmup)
                splan.set_profiling(False)
                host_ms, hc = time_host(splan, sa, sb, 5)
            others[label] = {"n": sn, "limbs": sl, "ms_device": round(dms, 4),
                             "e2e_host_ms": round(float(np.median(host_ms)), 3),
                             "e2e_matches_device": bool(np.array_equal(hc, dc))}
        result["other_configs"] = others

    if not args.no_batch:
        # one untimed small product first: HIP and kernel-module start-up belong to neither leg
        warm = workloads.make_products(1, 1, [4096], (16, 17))
        plain_products(polymul, warm)
        batched, products = batched_leg(polymul, workloads, args.batch_count, args.lanes)
        result["batched"] = batched
        if args.batch_count == 64:  # the golden digests cover the 64-product workloads
            result["batched_device"] = device_batch_leg(polymul, torch, products, args.lanes)
            # the mix of small products (N <= 8192, L <= 128), none of which fills the GPU alone
            result["batched_device_small"] = device_batch_leg(polymul, torch, workloads.batch_small_products(),
                                                              args.lanes, golden_key="batch_small")
        del products
        result["resident_operand"] = resident_leg(polymul, workloads, 64)
        capacity = batched["batched_products_per_s"]
        result["served"] = served_leg(polymul, workloads, args.serve_count, args.lanes,
                                      [0.3 * capacity, 0.6 * capacity, 0.9 * capacity])
        print("[bench] batched/served legs done", file=sys.stderr, flush=True)

    if not args.no_cpu:
        cpu_s, scale, exact, reference = cpu_baseline(q, a, b, polymul.Plan, args.cpu_sample)
        result["cpu_baseline"] = {"value": round(cpu_s * scale * 1000.0, 1), "unit": "ms/product", "cores": 1,
                                  "kind": "port", "extrapolated": True,
                                  "sample": "GMP Kronecker product (oracle, 1 core) of the first %d coefficients "
                                            "as an N=%d product with the same q: %.2f s, scaled to N=%d by "
                                            "B*log2(B) of the packed integer sizes (x%.2f)"
                                            % (args.cpu_sample, args.cpu_sample, cpu_s, n, scale),
                                  "sample_gpu_exact": exact}
        # full-size oracle products, timed: N=1024 (256-bit q) in this process,
        # N=32768 (27,776-bit q) alone in a child process, then one such product
        # per core on every core this process may use (the CPU's throughput)
        sn, sw, sl, sq, sa, sb = workloads.seeded_case("small")
        t0 = time.perf_counter()
        reference.kronecker_gmp(sa, sb, sq)
        small_ms = (time.perf_counter() - t0) * 1000.0
        mid_alone, _ = _oracle_processes("mid", 1)
        cores = len(os.sched_getaffinity(0))
        workers = min(cores, 16)
        mid_each, mid_wall = _oracle_processes("mid", workers)
        timed = {"B-small_ms": round(small_ms, 2), "B-mid_ms": round(mid_alone[0] * 1000.0, 1)}
        result["cpu_baseline"]["timed_full_products"] = timed
        gpu_mid = result.get("other_configs", {}).get("B-mid", {})
        result["cpu_baseline"]["throughput"] = {
            "config": "B-mid (N=32768, L=868), one oracle product per core, all started together",
            "processes": workers, "cores_available": cores, "cpu_model": cpu_model(),
            "products_per_s": round(workers / mid_wall, 3),
            "mean_product_s": round(float(np.mean(mid_each)), 2),
            "gpu_products_per_s_device": round(1000.0 / gpu_mid["ms_device"], 1) if gpu_mid else None,
            "gpu_products_per_s_host_vectors": round(1000.0 / gpu_mid["e2e_host_ms"], 1) if gpu_mid else None}
        # sampled bit-exactness of the timed full-size product (checker only)
        checks = []
        for i in (0, n - 1):
            checks.append(bool(np.array_equal(c[i], reference.coefficient_gmp(a, b, q, i))))
        result["exact_sampled_coefficients"] = all(checks)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
