"""End-to-end through the drop-in surface: point directories written the way
the reference's tools/validate_local.py does (tests/reference.py:46-75), run by
the fherma_init/fherma_run/fherma_free case runner, checked byte-for-byte
(tests/reference.py:78-90)."""
import json
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
SOLUTION = ROOT / "fherma-polynomial-cuda_amd" / "solution"


@pytest.mark.parametrize("n,width", [(64, 33), (1024, 512), (32768, 868)])
def test_case_runner_matches_oracle(oracle, tmp_path, n, width):
    print("[gpu-case] driver", n, width, flush=True)
    point = tmp_path / ("n%d_w%d" % (n, width))
    info = oracle.save_point(point, n, width, seeds=(101, 202), edge_cases=True)
    proc = subprocess.run([str(SOLUTION), str(point)], capture_output=True, text=True, timeout=300)
    assert proc.returncode == 0, proc.stderr
    reports = oracle.verify_point(point)
    assert len(reports) == info["cases"] and all(r["passed"] for r in reports)
    results = json.loads((point / "out" / "results.json").read_text())
    assert results["init_s"] > 0
    assert [c["status"] for c in results["cases"]] == ["ok"] * info["cases"]


def test_case_runner_records_bad_input_as_crashed(oracle, tmp_path):
    point = tmp_path / "bad"
    oracle.save_point(point, 64, 33, seeds=(1,), edge_cases=False)
    (point / "cases" / "000000" / "a.bin").write_bytes(b"\0" * 12)  # wrong size
    proc = subprocess.run([str(SOLUTION), str(point)], capture_output=True, text=True, timeout=300)
    assert proc.returncode == 0, proc.stderr
    results = json.loads((point / "out" / "results.json").read_text())
    assert results["cases"][0]["status"] == "crashed"
    assert "reading the case" in results["cases"][0]["note"]
