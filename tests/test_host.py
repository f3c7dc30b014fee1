"""Host-side machinery of the library checked on the CPU, built from its own
sources with g++ (no GPU, no HIP calls)."""
import shutil
import subprocess
from pathlib import Path

import pytest


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ to build the pool check")
def test_copy_pool_under_stress(tmp_path):
    """The host copy workers' hand-offs (spin, sleep, wake by job or by product
    start) never lose a job or a byte (tests/pool_check.cpp)."""
    root = Path(__file__).resolve().parent.parent
    csrc = root / "fherma-polynomial-cuda_amd" / "csrc"
    exe = tmp_path / "pool_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(csrc), str(root / "tests" / "pool_check.cpp"),
                    "-o", str(exe), "-lpthread"], check=True, capture_output=True, timeout=300)
    proc = subprocess.run([str(exe), "1500"], capture_output=True, text=True, timeout=600)
    assert proc.returncode == 0, proc.stdout[-2000:]
    assert proc.stdout.startswith("ok ")


@pytest.mark.skipif(shutil.which("g++") is None or not Path("/opt/rocm/include/hip/hip_runtime.h").is_file(),
                    reason="needs g++ and the HIP headers to build the arena check")
def test_device_arena_with_a_fake_device(tmp_path):
    """The server's device arena (chunks on demand, admission, coalescing, idle
    chunks given back, refusals) driven by a fake allocator on the CPU
    (tests/arena_check.cpp)."""
    root = Path(__file__).resolve().parent.parent
    csrc = root / "fherma-polynomial-cuda_amd" / "csrc"
    exe = tmp_path / "arena_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-I", str(csrc),
                    str(root / "tests" / "arena_check.cpp"), "-o", str(exe), "-lpthread"],
                   check=True, capture_output=True, timeout=300)
    proc = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert proc.returncode == 0, proc.stdout[-2000:] + proc.stderr[-2000:]
    assert proc.stdout.strip() == "ok"
