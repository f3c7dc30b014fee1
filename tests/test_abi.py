"""The C ABI library loads on a machine without a GPU and exports every entry
point the headers under include/ declare; the fherma shim exports the three
harness functions; the case runner starts and reports usage.  No compute
calls are made here."""
import ctypes
import re
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "fherma-polynomial-cuda_amd"


def _declared_c_functions():
    text = (ROOT / "include" / "polymul.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:pm_status|void|const char\s*\*)\s*(pm_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_function(polymul):
    lib = ctypes.CDLL(str(polymul.LIBRARY_PATH))
    names = _declared_c_functions()
    assert len(names) >= 12
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_abi(polymul):
    lib = polymul.load_library()
    for name in _declared_c_functions():
        assert getattr(lib, name).argtypes is not None, name  # the binding declared its signature


def test_fherma_shim_exports_the_harness_entry_points():
    out = subprocess.run(["nm", "-D", "--defined-only", str(PKG / "libfherma_amd.so")], capture_output=True,
                         text=True, check=True).stdout
    demangled = subprocess.run(["c++filt"], input=out, capture_output=True, text=True, check=True).stdout
    for signature in ("fherma_init(fherma::Point const&)", "fherma_run(void*, fherma::Inputs const&)",
                      "fherma_free(void*)"):
        assert signature in demangled, signature


def test_case_runner_usage():
    proc = subprocess.run([str(PKG / "solution")], capture_output=True, text=True)
    assert proc.returncode == 2 and "usage" in proc.stderr


def test_last_error_is_thread_local_and_readable(polymul):
    lib = polymul.load_library()
    handle = ctypes.c_void_p()
    status = lib.pm_plan_create_host(3, 7, 1, (ctypes.c_uint32 * 1)(97), ctypes.byref(handle))
    assert status == 1
    assert b"power of two" in lib.pm_last_error()


def test_one_hip_runtime_in_either_import_order(tmp_path):
    """Loading the library before or after `import torch` leaves one HIP runtime
    mapped (PyTorch's wheel ships its own copy under the unversioned name; a
    second copy in the process would see no GPU)."""
    import sys

    script = tmp_path / "order.py"
    script.write_text(
        "import importlib.util, sys\n"
        "spec = importlib.util.spec_from_file_location('pm_polymul', %r)\n"
        "m = importlib.util.module_from_spec(spec); spec.loader.exec_module(m)\n"
        "if sys.argv[1] == 'lib-first':\n    m.load_library(); import torch\n"
        "else:\n    import torch; m.load_library()\n"
        "print(len(m.hip_runtimes()))\n" % str(PKG / "polymul.py"))
    for order in ("lib-first", "torch-first"):
        proc = subprocess.run([sys.executable, str(script), order], capture_output=True, text=True, timeout=600)
        assert proc.returncode == 0, proc.stderr[-2000:]
        assert proc.stdout.strip() == "1", (order, proc.stdout)
