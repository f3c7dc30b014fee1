"""Builds the gfx950 library and its C++ front ends in-tree with hipcc.

  libpolymul_amd.so     kernels.hip + api.cpp + serve.cpp + plan.cpp   (C ABI, include/polymul.h)
  libfherma_amd.so      fherma_shim.cpp linked to the above (include/fherma.h)
  solution              driver_main.cpp: the main.cpp-compatible case runner
  libpm_calibrate.so    calibrate.hip: register-only INT8 MFMA and 32-bit multiply
                        throughput kernels (the measured peaks beside the datasheet ones)

Every translation unit is compiled for gfx950 only (hipcc treats the .cpp
sources as HIP too), with packed-FP32 instructions disabled (the kernels are
integer-only; the flag is a safety net).
"""
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I", str(INCLUDE), "-I", str(CSRC)]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

LIB = PKG / "libpolymul_amd.so"
CALIBRATE_LIB = PKG / "libpm_calibrate.so"
FHERMA_LIB = PKG / "libfherma_amd.so"
SOLUTION = PKG / "solution"


def _run(cmd, verbose):
    if verbose:
        print(" ".join(str(c) for c in cmd), flush=True)
    proc = subprocess.run([str(c) for c in cmd], capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("build failed: %s\n%s%s" % (" ".join(str(c) for c in cmd), proc.stdout, proc.stderr))
    return proc


def build(verbose=False):
    obj_dir = PKG / "_build"
    obj_dir.mkdir(exist_ok=True)
    objects = []
    for src in ("kernels.hip", "api.cpp", "serve.cpp", "plan.cpp"):
        obj = obj_dir / (src.rsplit(".", 1)[0] + ".o")
        flags = COMMON + DEVICE  # .cpp sources hold no kernels; gfx950 keeps hipcc off its default target
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC] + flags + lang + ["-c", CSRC / src, "-o", obj], verbose)
        objects.append(obj)
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH] + objects + ["-o", LIB], verbose)
    shim = CSRC / "fherma_shim.cpp"
    if shim.is_file():
        _run([HIPCC] + COMMON + DEVICE + ["-shared", shim, "-L", PKG, "-lpolymul_amd",
                                 "-Wl,-rpath,$ORIGIN", "-o", FHERMA_LIB], verbose)
        driver = CSRC / "driver_main.cpp"
        if driver.is_file():
            _run([HIPCC] + COMMON + DEVICE + [driver, "-L", PKG, "-lfherma_amd", "-lpolymul_amd",
                                     "-Wl,-rpath,$ORIGIN", "-o", SOLUTION], verbose)
    calibrate = CSRC / "calibrate.hip"
    if calibrate.is_file():
        _run([HIPCC] + COMMON + DEVICE + ["-x", "hip", "-shared", calibrate, "-o", CALIBRATE_LIB], verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
