"""TEST INFRASTRUCTURE ONLY: builds the CPU oracle (and, where the reference tree
is present, the reference's own host helpers) as shared libraries.

  oracle/liboracle.so              oracle/kronecker.c linked against libgmp.so.10
  oracle/_ref/libref_big_unsigned.so
                                   candidates/ordered-streaming/big_unsigned.h,
                                   compiled from the reference tree where it
                                   lies, wrapped by oracle/ref_big_unsigned_shim.cpp

Nothing from the reference is copied: the header is included by path.  The
_ref library is only built in a container that has /root/reference; GPU boxes
use the file built here.
"""
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REFERENCE = Path("/root/reference")
REF_HEADER_DIR = REFERENCE / "candidates" / "ordered-streaming"


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("oracle build failed: %s\n%s%s" % (" ".join(cmd), proc.stdout, proc.stderr))


def build_oracle(verbose=False):
    out = HERE / "liboracle.so"
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-Wall", "-Wextra",
           str(HERE / "kronecker.c"), "-o", str(out), "-l:libgmp.so.10"]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    return out


def build_reference_helpers(verbose=False):
    """Returns the built path, or None when the reference tree is absent."""
    if not (REF_HEADER_DIR / "big_unsigned.h").is_file():
        return None
    out_dir = HERE / "_ref"
    out_dir.mkdir(exist_ok=True)
    out = out_dir / "libref_big_unsigned.so"
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-I", str(REF_HEADER_DIR),
           str(HERE / "ref_big_unsigned_shim.cpp"), "-o", str(out)]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    return out


def build(verbose=False):
    build_oracle(verbose)
    build_reference_helpers(verbose)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
