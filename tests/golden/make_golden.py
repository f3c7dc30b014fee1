"""Writes the committed fixtures under tests/golden/ (run from the repo root:
`python tests/golden/make_golden.py`).

  points.json      q of the reference's competition and validation points:
                   the NTT-friendly prime below 2^W with q = 1 (mod 2N),
                   tests/reference.py:9-16, via the oracle's modulus().
  ref_primes.json  the prime lists the reference recorded for its two
                   competition points (diagnostics/tensor-quotient/ARITHMETIC.json,
                   copied as data: N, field count and the primes).
  small_*.npz      the small plain configuration (N=1024, L=8, random odd
                   256-bit q) with its edge cases and the oracle's outputs
                   (GMP Kronecker, cross-checked with Python-int Kronecker).
"""
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
HERE = Path(__file__).resolve().parent

from oracle import reference  # noqa: E402

VALIDATE_POINTS = [(1, 2), (2, 5), (8, 16), (64, 31), (64, 33), (128, 128), (256, 256),
                   (1024, 512), (8192, 256), (32768, 868), (65536, 1747)]
STRESS_POINTS = [(128, 4096), (131072, 128)]


def write_points():
    points = {}
    for n, width in VALIDATE_POINTS + STRESS_POINTS:
        q = reference.modulus(n, width)
        points["%d/%d" % (n, width)] = {"N": n, "W": width, "L": (width + 31) // 32, "q": hex(q),
                                        "gap": (1 << width) - q}
    (HERE / "points.json").write_text(json.dumps(points, indent=1) + "\n")


def write_ref_primes():
    src = Path("/root/reference/diagnostics/tensor-quotient/ARITHMETIC.json")
    data = json.loads(src.read_text())
    out = [{"N": rep["N"], "fields": rep["fields"], "primes": rep["primes"]} for rep in data["reports"]]
    (HERE / "ref_primes.json").write_text(json.dumps({"source": "diagnostics/tensor-quotient/ARITHMETIC.json",
                                                      "reports": out}, indent=1) + "\n")


def write_small():
    import importlib.util
    spec = importlib.util.spec_from_file_location("workloads", ROOT / "fherma-polynomial-cuda_amd" / "workloads.py")
    workloads = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(workloads)
    n, width, limbs, q, a, b = workloads.seeded_case("small")
    qi = reference.words_to_int(q)
    cases = {"random": (a, b)}
    z = np.zeros((n, limbs), dtype=np.uint32)
    unit = z.copy()
    unit[0, 0] = 1
    xk = z.copy()
    xk[n - 1, 0] = 1
    qm1 = np.tile(workloads.q_minus_one(q), (n, 1))
    cases.update({"zero": (z, b), "unit": (a, unit), "max": (qm1, qm1), "impulse_top": (xk, a)})
    arrays = {"q": q}
    for name, (x, y) in cases.items():
        c = reference.kronecker_gmp(x, y, q)
        assert reference.from_limbs(c) == reference.kronecker(reference.from_limbs(x), reference.from_limbs(y), qi)
        arrays[name + "_a"] = x
        arrays[name + "_b"] = y
        arrays[name + "_c"] = c
    np.savez_compressed(HERE / "small_cases.npz", **arrays)


if __name__ == "__main__":
    write_points()
    write_ref_primes()
    write_small()
    print("fixtures written to", HERE)
