"""TEST INFRASTRUCTURE ONLY — the CPU oracle for c = a*b mod (X^N + 1, q).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import
this module, and only as the checker; the product library never calls it.

Every function restates a piece of the reference (file:line cited):

  modulus            tests/reference.py:9-16    NTT-friendly prime q < 2^W
  schoolbook         tests/reference.py:19-26   O(N^2) signed negacyclic sum
  kronecker          tests/reference.py:29-43   Kronecker packing (Python int)
  kronecker_gmp      tests/reference.py:29-43   same algorithm through GMP (C)
  save_point         tests/reference.py:46-75   case directory writer
  verify_point       tests/reference.py:78-90   byte-exact case checker
  reference_primes   candidates/ordered-streaming/solve.cu:905-913 prime walk
  crt_reconstruct    candidates/pipelined-direct-crt/EXACTNESS.md:1-57 and
                     tests/direct_crt_math.py:9-31 (direct CRT, fixed-point
                     quotient, shortened sum T = sum v_k M_k + K R)
  montgomery_redc    candidates/gpu-refined-columns/TENSOR_CONVERSION.md and
                     tests/tensor_montgomery_math.py:13-25 (subtractive REDC)

How each is pinned (tests/test_oracle.py): kronecker == schoolbook for N <= 64
(the reference's own cross-check, tests/reference.py:70-71); kronecker (Python
int) == kronecker_gmp (GMP) on every fixture, two independent big-integer
libraries; reference_primes == the 57/114-prime lists recorded in
diagnostics/tensor-quotient/ARITHMETIC.json and == the reference's own compiled
helpers (oracle/_ref); modulus uses mpz_probab_prime_p, the routine
gmpy2.is_prime wraps, and is cross-checked against sympy.isprime; the CRT and
REDC restatements are checked by brute force against Python's pow().
"""
import ctypes
import json
import random
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

# ----------------------------------------------------------------------------
# shared-library handles
# ----------------------------------------------------------------------------
_lib = None
_ref = None


def library():
    """oracle/liboracle.so (built by oracle/build_oracle.py)."""
    global _lib
    if _lib is None:
        path = HERE / "liboracle.so"
        if not path.is_file():
            from . import build_oracle  # noqa: WPS433 (lazy: build on first use)
            build_oracle.build_oracle()
        lib = ctypes.CDLL(str(path))
        u32p = ctypes.POINTER(ctypes.c_uint32)
        for name, args in {
            "oracle_kronecker": [ctypes.c_uint32, ctypes.c_uint32, u32p, u32p, u32p, u32p],
            "oracle_schoolbook": [ctypes.c_uint32, ctypes.c_uint32, u32p, u32p, u32p, u32p],
            "oracle_coefficient": [ctypes.c_uint32, ctypes.c_uint32, u32p, u32p, u32p,
                                   ctypes.c_uint32, u32p],
            "oracle_signed_convolution": [ctypes.c_uint32, ctypes.c_uint32, u32p, u32p,
                                          ctypes.c_uint32, u32p, ctypes.c_uint32,
                                          ctypes.POINTER(ctypes.c_int32)],
            "oracle_probab_prime": [ctypes.c_uint32, u32p, ctypes.c_int],
        }.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _lib = lib
    return _lib


def reference_helpers():
    """oracle/_ref/libref_big_unsigned.so, or None where it was not built."""
    global _ref
    if _ref is None:
        path = HERE / "_ref" / "libref_big_unsigned.so"
        if not path.is_file():
            return None
        lib = ctypes.CDLL(str(path))
        u32p = ctypes.POINTER(ctypes.c_uint32)
        lib.ref_is_prime.argtypes = [ctypes.c_uint32]
        lib.ref_is_prime.restype = ctypes.c_int
        lib.ref_modulo.argtypes = [u32p, ctypes.c_size_t, u32p, ctypes.c_size_t, u32p, ctypes.c_size_t]
        lib.ref_modulo.restype = ctypes.c_int
        lib.ref_modulo_word.argtypes = [u32p, ctypes.c_size_t, ctypes.c_uint32]
        lib.ref_modulo_word.restype = ctypes.c_uint32
        lib.ref_divide_word.argtypes = [u32p, ctypes.c_size_t, ctypes.c_uint32, u32p, ctypes.c_size_t]
        lib.ref_divide_word.restype = ctypes.c_int
        lib.ref_multiply.argtypes = [u32p, ctypes.c_size_t, u32p, ctypes.c_size_t, u32p, ctypes.c_size_t]
        lib.ref_multiply.restype = ctypes.c_int
        lib.ref_prime_walk.argtypes = [ctypes.c_uint32, u32p, ctypes.c_size_t, u32p, ctypes.c_size_t]
        lib.ref_prime_walk.restype = ctypes.c_int
        _ref = lib
    return _ref


# ----------------------------------------------------------------------------
# limb helpers
# ----------------------------------------------------------------------------
def int_to_words(value, words):
    """Non-negative int -> little-endian uint32 array of `words` limbs."""
    return np.frombuffer(int(value).to_bytes(4 * words, "little"), dtype="<u4").astype(np.uint32)


def words_to_int(array):
    return int.from_bytes(np.ascontiguousarray(array, dtype="<u4").tobytes(), "little")


def to_limbs(values, limbs):
    """List of ints -> [N, L] uint32 (coefficient-major, little-endian limbs)."""
    raw = b"".join(int(x).to_bytes(4 * limbs, "little") for x in values)
    return np.frombuffer(raw, dtype="<u4").astype(np.uint32).reshape(len(values), limbs)


def from_limbs(array):
    array = np.ascontiguousarray(array, dtype="<u4")
    n, limbs = array.shape
    raw = array.tobytes()
    return [int.from_bytes(raw[i * 4 * limbs:(i + 1) * 4 * limbs], "little") for i in range(n)]


def _ptr(array):
    return array.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


# ----------------------------------------------------------------------------
# primality (REF:12 uses gmpy2.is_prime(q, 40) == mpz_probab_prime_p(q, 40))
# ----------------------------------------------------------------------------
def is_probable_prime(x, reps=40):
    x = int(x)
    if x < 2:
        return False
    words = max(1, (x.bit_length() + 31) // 32)
    buf = int_to_words(x, words)
    return library().oracle_probab_prime(words, _ptr(buf), reps) > 0


def modulus(n, width):
    """Largest prime q < 2^width with q = 1 (mod 2n): tests/reference.py:9-16."""
    step = 2 * n
    q = ((1 << width) - 2) // step * step + 1
    while q > 1 and not is_probable_prime(q, 40):
        q -= step
    if q <= 1:
        raise ValueError("No modulus exists for this point")
    return q


# ----------------------------------------------------------------------------
# the end-to-end oracle
# ----------------------------------------------------------------------------
def schoolbook(a, b, q):
    """tests/reference.py:19-26 (pure Python; for small N)."""
    n = len(a)
    out = [0] * n
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            k = i + j
            out[k % n] += x * y if k < n else -x * y
    return [x % q for x in out]


def kronecker_stride(n, q):
    return max(1, ((n * (q - 1) ** 2).bit_length() + 7) // 8)


def kronecker(a, b, q):
    """tests/reference.py:29-43 with Python's own big integers (no GMP)."""
    n = len(a)
    stride = kronecker_stride(n, q)
    left = int.from_bytes(b"".join(x.to_bytes(stride, "little") for x in a), "little")
    right = int.from_bytes(b"".join(x.to_bytes(stride, "little") for x in b), "little")
    raw = (left * right).to_bytes(2 * n * stride, "little")
    return [(int.from_bytes(raw[i * stride:(i + 1) * stride], "little")
             - int.from_bytes(raw[(i + n) * stride:(i + n + 1) * stride], "little")) % q
            for i in range(n)]


def signed_convolution(a, b):
    """Exact integer negacyclic coefficients x_i (before mod q), via Kronecker
    packing with one extra sign-free slot: x_i = lo_i - hi_i (REF:40-43)."""
    n = len(a)
    bound = max(1, n * max(max(a, default=0), 1) * max(max(b, default=0), 1))
    stride = (bound.bit_length() + 8) // 8
    left = int.from_bytes(b"".join(x.to_bytes(stride, "little") for x in a), "little")
    right = int.from_bytes(b"".join(x.to_bytes(stride, "little") for x in b), "little")
    raw = (left * right).to_bytes(2 * n * stride, "little")
    return [int.from_bytes(raw[i * stride:(i + 1) * stride], "little")
            - int.from_bytes(raw[(i + n) * stride:(i + n + 1) * stride], "little")
            for i in range(n)]


def kronecker_gmp(a, b, q_words):
    """tests/reference.py:29-43 through GMP: a, b are [N, L] uint32 arrays."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    q_words = np.ascontiguousarray(q_words, dtype=np.uint32)
    n, limbs = a.shape
    c = np.zeros_like(a)
    rc = library().oracle_kronecker(n, limbs, _ptr(q_words), _ptr(a), _ptr(b), _ptr(c))
    if rc:
        raise RuntimeError("oracle_kronecker failed: %d" % rc)
    return c


def coefficient_gmp(a, b, q_words, i):
    """One output coefficient by the direct signed sum (REF:19-26), via GMP."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    q_words = np.ascontiguousarray(q_words, dtype=np.uint32)
    n, limbs = a.shape
    out = np.zeros(limbs, dtype=np.uint32)
    rc = library().oracle_coefficient(n, limbs, _ptr(q_words), _ptr(a), _ptr(b), int(i), _ptr(out))
    if rc:
        raise RuntimeError("oracle_coefficient failed: %d" % rc)
    return out


def schoolbook_gmp(a, b, q_words):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    q_words = np.ascontiguousarray(q_words, dtype=np.uint32)
    n, limbs = a.shape
    c = np.zeros_like(a)
    rc = library().oracle_schoolbook(n, limbs, _ptr(q_words), _ptr(a), _ptr(b), _ptr(c))
    if rc:
        raise RuntimeError("oracle_schoolbook failed: %d" % rc)
    return c


# ----------------------------------------------------------------------------
# case directories (tests/reference.py:46-90)
# ----------------------------------------------------------------------------
def edge_pairs(n, q):
    """The four fixed cases of tests/reference.py:52-57."""
    zeros, ones, maximum = [0] * n, [1] * n, [q - 1] * n
    impulse, shifted = zeros.copy(), zeros.copy()
    impulse[0] = 1
    shifted[-1] = q - 1
    return [(zeros, maximum), (ones, maximum), (maximum, maximum), (impulse, shifted)]


def random_pair(n, q, seed):
    """tests/reference.py:58-60: the reference's exact random stream."""
    rng = random.Random(seed)
    return [rng.randrange(q) for _ in range(n)], [rng.randrange(q) for _ in range(n)]


def stress_pairs(n, q, limbs):
    """The boundary pairs of tools/stress_local.py:23-27."""
    last = [0] * n
    last[-1] = q - 1
    ones = [1] * n
    alternating = [q - 1 if i % 2 else 0 for i in range(n)]
    limb_boundaries = [min(q - 1, (1 << (32 * (i % limbs + 1))) - 1) for i in range(n)]
    return [(last, last), (alternating, ones), (limb_boundaries, limb_boundaries),
            ([q // 2] * n, [q // 2 + 1] * n)]


def save_point(path, n, width, q=None, seeds=(1, 2), edge_cases=True, extra_pairs=(), expected=True):
    """tests/reference.py:46-75 (optionally with a caller-chosen q)."""
    path = Path(path)
    q = modulus(n, width) if q is None else int(q)
    limbs = (width + 31) // 32
    byte_width = 4 * limbs
    pairs = list(edge_pairs(n, q)) if edge_cases else []
    for seed in seeds:
        pairs.append(random_pair(n, q, seed))
    pairs.extend(extra_pairs)
    (path / "point").mkdir(parents=True, exist_ok=True)
    (path / "point/q.bin").write_bytes(q.to_bytes(byte_width, "little"))
    (path / "manifest.json").write_text(json.dumps({"point": {"N": n, "W": width, "L": limbs},
                                                    "cases": len(pairs)}))
    q_words = int_to_words(q, limbs)
    for index, (a, b) in enumerate(pairs):
        directory = path / "cases" / ("%06d" % index)
        directory.mkdir(parents=True, exist_ok=True)
        for name, values in (("a", a), ("b", b)):
            (directory / ("%s.bin" % name)).write_bytes(b"".join(x.to_bytes(byte_width, "little")
                                                                 for x in values))
        if expected:
            c = kronecker_gmp(to_limbs(a, limbs), to_limbs(b, limbs), q_words)
            if n <= 64:
                assert from_limbs(c) == schoolbook(a, b, q)
            reference = path / "reference" / ("%06d" % index)
            reference.mkdir(parents=True, exist_ok=True)
            (reference / "c.bin").write_bytes(c.astype("<u4").tobytes())
    return {"N": n, "W": width, "L": limbs, "cases": len(pairs), "q_gap": (1 << width) - q}


def verify_point(path):
    """tests/reference.py:78-90: byte-exact comparison of out/ with reference/."""
    path = Path(path)
    manifest = json.loads((path / "manifest.json").read_text())
    reports = []
    for index in range(manifest["cases"]):
        expected = (path / "reference" / ("%06d" % index) / "c.bin").read_bytes()
        actual = (path / "out" / ("%06d" % index) / "c.bin").read_bytes()
        passed = expected == actual
        reports.append({"case": index, "passed": passed})
        if not passed:
            word = next((i for i in range(min(len(expected), len(actual))) if expected[i] != actual[i]), None)
            raise AssertionError("%s: case %d differs at byte %s; lengths %d / %d"
                                 % (path, index, word, len(actual), len(expected)))
    return reports


# ----------------------------------------------------------------------------
# host-plan restatements
# ----------------------------------------------------------------------------
def is_prime_u32(p):
    """Deterministic Miller-Rabin, witnesses {2,3,5,7,11}
    (candidates/ordered-streaming/big_unsigned.h:83-97)."""
    if p < 2:
        return False
    for a in (2, 3, 5, 7, 11):
        if p == a:
            return True
        if p % a == 0:
            return False
    d, shift = p - 1, 0
    while not d & 1:
        d >>= 1
        shift += 1
    for a in (2, 3, 5, 7, 11):
        x = pow(a, d, p)
        if x in (1, p - 1):
            continue
        for _ in range(1, shift):
            x = x * x % p
            if x == p - 1:
                break
        else:
            return False
    return True


def reference_primes(n, q, step=None):
    """fherma_init's walk (candidates/ordered-streaming/solve.cu:905-913):
    descending p = 1 (mod step) from floor((2^31-2)/step)*step+1, primes only,
    until prod(p) > 4*N*(q-1)^2; ValueError below 2^30.  `step` defaults to 2N."""
    step = 2 * n if step is None else step
    bound = 4 * n * (q - 1) ** 2
    product, primes = 1, []
    candidate = ((1 << 31) - 2) // step * step + 1
    while product <= bound:
        if candidate < (1 << 30):
            raise ValueError("insufficient independent NTT fields")
        if is_prime_u32(candidate):
            primes.append(candidate)
            product *= candidate
        candidate -= step
    return primes


# ----------------------------------------------------------------------------
# arithmetic identities (restated, checked by brute force in the tests)
# ----------------------------------------------------------------------------
R32 = 1 << 32


def montgomery_redc(value, p):
    """Subtractive REDC of an integer value (any sign, |value| < 2^63):
    (value - m*p)/2^32 with m = (value mod 2^32)*p^-1 mod 2^32, then brought into
    [0, p).  TENSOR_CONVERSION.md:9-20; tests/tensor_montgomery_math.py:13-19."""
    inverse = pow(p, -1, R32)
    multiplier = ((value % R32) * inverse) % R32
    reduced = (value - multiplier * p) // R32
    return reduced % p


def crt_reconstruct(x, primes, q):
    """EXACTNESS.md:1-52 / tests/direct_crt_math.py:14-30 for one signed x with
    |x| < prod(primes)/4: returns (v, K, T) where v_k = x*P_k^-1 mod p_k,
    K = round(sum v_k/p_k) (the integer CRT quotient), and
    T = sum v_k*(P_k mod q) + K*((-P) mod q), with T = x (mod q)."""
    product = 1
    for p in primes:
        product *= p
    partial = [product // p for p in primes]
    v = [((x % p) * pow(m % p, -1, p)) % p for p, m in zip(primes, partial)]
    quotient = (sum(a * m for a, m in zip(v, partial)) - x) // product
    t = sum(a * (m % q) for a, m in zip(v, partial)) + quotient * ((-product) % q)
    return v, quotient, t


def fixed_point_quotient(v, primes):
    """EXACTNESS.md:22-35: K = (sum((v_k*floor((2^64-1)/p_k)) >> 32) + 2^31) >> 32."""
    total = sum((a * (((1 << 64) - 1) // p)) >> 32 for a, p in zip(v, primes))
    return (total + (1 << 31)) >> 32
