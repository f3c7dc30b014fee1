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


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ to build the pool check")
def test_copy_pool_under_thread_sanitizer(tmp_path):
    """The same stress run, shorter, built with ThreadSanitizer: posted jobs,
    the piece counter, sleeping and spinning workers and the destructor's
    drain show no data race."""
    root = Path(__file__).resolve().parent.parent
    csrc = root / "fherma-polynomial-cuda_amd" / "csrc"
    exe = tmp_path / "pool_check_tsan"
    build = subprocess.run(["g++", "-O1", "-g", "-fsanitize=thread", "-std=c++17", "-I", str(csrc),
                            str(root / "tests" / "pool_check.cpp"), "-o", str(exe), "-lpthread"],
                           capture_output=True, text=True, timeout=300)
    if build.returncode != 0:
        pytest.skip("no ThreadSanitizer runtime: " + build.stderr[-300:])
    proc = subprocess.run([str(exe), "120"], capture_output=True, text=True, timeout=600)
    assert "ThreadSanitizer" not in proc.stderr, proc.stderr[-4000:]
    assert proc.returncode == 0, proc.stdout[-2000:] + proc.stderr[-2000:]
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


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ to build the chunk check")
def test_pipeline_chunks_stay_inside_the_tiles(tmp_path):
    """Every chunk of the host pipeline is a non-empty set of tiles inside
    [0, ntiles), the chunks cover every tile and every coefficient once, and
    with column strips each strip's column stages are issued once, after its
    last piece is in and before its first piece goes out -- for every
    power-of-two N up to 2^24 and L up to 8192 limbs (tests/chunk_check.cpp).
    A chunk starting past the last tile used to make the packing kernel write
    past the packed-operand buffer (e.g. N=2048, L=4096: 6 chunks of 2 tiles
    over 8 tiles)."""
    root = Path(__file__).resolve().parent.parent
    csrc = root / "fherma-polynomial-cuda_amd" / "csrc"
    exe = tmp_path / "chunk_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(csrc), str(root / "tests" / "chunk_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True, timeout=300)
    proc = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert proc.returncode == 0, proc.stdout[-2000:]
    assert proc.stdout.startswith("ok ")


@pytest.mark.skipif(shutil.which("cmake") is None or not Path("/opt/rocm/bin/hipcc").is_file(),
                    reason="needs cmake and hipcc")
def test_harness_cmake_build(tmp_path):
    """The FHERMA runner's own build command (src/fherma.toml:11-20) on a copy
    of the sources: `cmake -B build -DCMAKE_BUILD_TYPE=Release && cmake --build
    build` must produce build/solution and the three harness entry points."""
    root = Path(__file__).resolve().parent.parent
    src = tmp_path / "solution_src"
    (src / "fherma-polynomial-cuda_amd").mkdir(parents=True)
    shutil.copy(root / "CMakeLists.txt", src)
    shutil.copytree(root / "include", src / "include")
    shutil.copytree(root / "fherma-polynomial-cuda_amd" / "csrc", src / "fherma-polynomial-cuda_amd" / "csrc")
    proc = subprocess.run("cmake -B build -DCMAKE_BUILD_TYPE=Release && cmake --build build -j 8", shell=True,
                          cwd=src, capture_output=True, text=True, timeout=900)
    assert proc.returncode == 0, proc.stdout[-3000:] + proc.stderr[-3000:]
    solution = src / "build" / "solution"
    assert solution.is_file()
    out = subprocess.run(["nm", "-D", "--defined-only", str(src / "build" / "libfherma_amd.so")],
                         capture_output=True, text=True, check=True).stdout
    names = subprocess.run(["c++filt"], input=out, capture_output=True, text=True, check=True).stdout
    for signature in ("fherma_init(fherma::Point const&)", "fherma_run(void*, fherma::Inputs const&)",
                      "fherma_free(void*)"):
        assert signature in names, signature
    usage = subprocess.run([str(solution)], capture_output=True, text=True)
    assert usage.returncode == 2 and "usage" in usage.stderr
