"""Seeded synthetic operands for the multiplier's workloads (numpy only).

Configurations (limb counts L; W = 32L bits of q unless stated):
  small   N=1024,  L=8     random odd 256-bit q, uniform operands
  mid     N=32768, L=868   random odd 27,776-bit q, uniform operands with 1% of
                           coefficients forced to q-1 and 1% to 0
  large   N=65536, L=1747  random odd 55,904-bit q, uniform operands
The reference's own competition points (q the NTT-friendly prime below 2^W,
W = 1747 or 868 bits) take q from tests/golden/points.json.

q is 32L random bits with the top bit and bit 0 set; coefficients are uniform
in [0, q) by rejection.  numpy's PCG64 with a fixed seed makes every case
reproducible on any machine with the same numpy.
"""
import numpy as np

CONFIGS = {
    "small": {"n": 1024, "limbs": 8, "seed": 0x5EED0001},
    "mid": {"n": 32768, "limbs": 868, "seed": 0x5EED0002},
    "large": {"n": 65536, "limbs": 1747, "seed": 0x5EED0003},
}


def words_to_int(words):
    return int.from_bytes(np.ascontiguousarray(words, dtype="<u4").tobytes(), "little")


def int_to_words(value, limbs):
    return np.frombuffer(int(value).to_bytes(4 * limbs, "little"), dtype="<u4").astype(np.uint32)


def random_modulus(limbs, rng):
    """32L random bits, top bit and bit 0 set."""
    q = rng.integers(0, 1 << 32, size=limbs, dtype=np.uint64).astype(np.uint32)
    q[-1] |= np.uint32(0x80000000)
    q[0] |= np.uint32(1)
    return q


def _ge_rows(values, q_words):
    """Boolean per row: value >= q (lexicographic from the top limb)."""
    rev = values[:, ::-1]
    qrev = q_words[::-1]
    diff = rev != qrev
    has = diff.any(axis=1)
    first = np.argmax(diff, axis=1)
    rows = np.arange(values.shape[0])
    return ~has | (rev[rows, first] > qrev[first])


def uniform_below(q_words, n, rng):
    """[n][L] uint32, each row uniform in [0, q)."""
    q_words = np.ascontiguousarray(q_words, dtype=np.uint32)
    limbs = q_words.size
    top = int(q_words[-1])
    if top == 0:
        raise ValueError("q's top limb must be non-zero")
    mask = np.uint32((1 << top.bit_length()) - 1)
    out = np.empty((n, limbs), dtype=np.uint32)
    pending = np.arange(n)
    while pending.size:
        draw = rng.integers(0, 1 << 32, size=(pending.size, limbs), dtype=np.uint64).astype(np.uint32)
        draw[:, -1] &= mask
        bad = _ge_rows(draw, q_words)
        good = ~bad
        out[pending[good]] = draw[good]
        pending = pending[bad]
    return out


def q_minus_one(q_words):
    q = np.array(q_words, dtype=np.uint32, copy=True)
    i = 0
    while q[i] == 0:
        q[i] = 0xFFFFFFFF
        i += 1
    q[i] -= np.uint32(1)
    return q


def seeded_case(name, seed=None):
    """(n, width_bits, limbs, q_words, a, b) for one of CONFIGS."""
    cfg = CONFIGS[name]
    n, limbs = cfg["n"], cfg["limbs"]
    rng = np.random.Generator(np.random.PCG64(cfg["seed"] if seed is None else seed))
    q = random_modulus(limbs, rng)
    a = uniform_below(q, n, rng)
    b = uniform_below(q, n, rng)
    if name == "mid":
        qm1 = q_minus_one(q)
        for arr in (a, b):
            idx = rng.permutation(n)
            k = n // 100
            arr[idx[:k]] = qm1
            arr[idx[k:2 * k]] = 0
    return n, 32 * limbs, limbs, q, a, b


def operands_for(n, q_words, seed):
    """Two uniform [n][L] operands below q from one seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    q_words = np.ascontiguousarray(q_words, dtype=np.uint32)
    return uniform_below(q_words, n, rng), uniform_below(q_words, n, rng)


def log_uniform_int(rng, lo, hi):
    import math

    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))


def make_products(count, seed, n_choices, limb_range):
    """Independent products [(n, limbs, q, a, b)], each with its own seeded
    random odd q and uniform operands."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(n_choices[rng.integers(0, len(n_choices))])
        limbs = log_uniform_int(rng, *limb_range)
        prng = np.random.default_rng([seed, i])
        q = random_modulus(limbs, prng)
        out.append((n, limbs, q, uniform_below(q, n, prng), uniform_below(q, n, prng)))
    return out


POWERS = [4096, 8192, 16384, 32768, 65536]
# The batched workload: 64 products, N in {2^12..2^16}, L log-uniform in
# [16, 1747] limbs, own q each.  The served workload: 48 requests, L in
# [32, 1747].  The resident-operand workload: 64 random a times one fixed b at
# N = 65536, L = 55.
BATCH = {"count": 64, "seed": 0x5EED0004, "limbs": (16, 1747)}
# A mix of small products only (N in {2^10..2^13}, L log-uniform in [8, 128]),
# none of which fills the GPU alone: what running products side by side gains
# when the device is not already saturated by the largest of them.
SMALL_POWERS = [1024, 2048, 4096, 8192]
BATCH_SMALL = {"count": 64, "seed": 0x5EED0007, "limbs": (8, 128)}
SERVED = {"count": 48, "seed": 0x5EED0006, "limbs": (32, 1747)}
RESIDENT = {"count": 64, "seed": 0x5EED0005, "n": 65536, "limbs": 55}


def batch_products(count=None):
    return make_products(BATCH["count"] if count is None else count, BATCH["seed"], POWERS, BATCH["limbs"])


def batch_small_products(count=None):
    return make_products(BATCH_SMALL["count"] if count is None else count, BATCH_SMALL["seed"], SMALL_POWERS,
                         BATCH_SMALL["limbs"])


def served_products(count=None):
    return make_products(SERVED["count"] if count is None else count, SERVED["seed"], POWERS, SERVED["limbs"])


def resident_case(count=None):
    """(n, limbs, q, b, [a_0 .. a_{count-1}])."""
    n, limbs = RESIDENT["n"], RESIDENT["limbs"]
    rng = np.random.default_rng(RESIDENT["seed"])
    q = random_modulus(limbs, rng)
    b = uniform_below(q, n, rng)
    a_list = [uniform_below(q, n, rng) for _ in range(RESIDENT["count"] if count is None else count)]
    return n, limbs, q, b, a_list


def digest(c):
    """SHA-256 of an [N][L] result as little-endian 32-bit words."""
    import hashlib

    return hashlib.sha256(np.ascontiguousarray(c, dtype="<u4").tobytes()).hexdigest()


def write_point_dir(root, n, width, limbs, q_words, cases):
    """A case directory in the harness's layout (src/main.cpp): manifest.json,
    point/q.bin, cases/NNNNNN/{a,b}.bin; `cases` is a list of (a, b)."""
    import json
    from pathlib import Path

    root = Path(root)
    (root / "point").mkdir(parents=True, exist_ok=True)
    (root / "manifest.json").write_text(json.dumps({"point": {"N": n, "W": width, "L": limbs}, "cases": len(cases)}))
    np.ascontiguousarray(q_words, dtype="<u4").tofile(root / "point" / "q.bin")
    for i, (a, b) in enumerate(cases):
        d = root / "cases" / ("%06d" % i)
        d.mkdir(parents=True, exist_ok=True)
        np.ascontiguousarray(a, dtype="<u4").tofile(d / "a.bin")
        np.ascontiguousarray(b, dtype="<u4").tofile(d / "b.bin")


def write_reference_point_dir(root, key="65536/1747", cases=21, seed0=7000):
    """The reference's competition point `key` (N/W, q from
    tests/golden/points.json) as a case directory, case i from seed seed0 + i."""
    import json
    from pathlib import Path

    points = json.loads((Path(__file__).resolve().parent.parent / "tests" / "golden" / "points.json").read_text())
    p = points[key]
    q = int_to_words(int(p["q"], 16), p["L"])
    write_point_dir(root, p["N"], p["W"], p["L"], q, [operands_for(p["N"], q, seed0 + i) for i in range(cases)])


if __name__ == "__main__":
    import argparse

    parser = argparse.ArgumentParser(description="write a reference competition point as a case directory")
    parser.add_argument("root")
    parser.add_argument("--point", default="65536/1747", help="N/W key of tests/golden/points.json")
    parser.add_argument("--cases", type=int, default=21)
    args = parser.parse_args()
    write_reference_point_dir(args.root, args.point, args.cases)
