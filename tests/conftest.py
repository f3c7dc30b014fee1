import ctypes
import importlib.util
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "fherma-polynomial-cuda_amd"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running case")


def _load(name):
    spec = importlib.util.spec_from_file_location("pm_" + name, PKG / (name + ".py"))
    module = importlib.util.module_from_spec(spec)
    sys.modules["pm_" + name] = module
    spec.loader.exec_module(module)
    return module


@pytest.fixture(scope="session")
def polymul():
    return _load("polymul")


@pytest.fixture(scope="session")
def workloads():
    return _load("workloads")


@pytest.fixture(scope="session")
def oracle():
    from oracle import reference
    return reference


class DeviceMemory:
    """Device buffers from the HIP runtime the polymul library is bound to.

    torch is not used here: its HIP libraries link the unversioned
    libamdhip64.so, so importing torch after the library has loaded ROCm's
    libamdhip64.so.7 brings a second HIP runtime into the process, and that
    one sees no GPU."""

    def __init__(self, polymul):
        polymul.load_library()
        self.rt = ctypes.CDLL("libamdhip64.so.7")  # the copy already loaded by the library
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []

    def _check(self, status, what):
        if status != 0:
            raise RuntimeError("%s failed with HIP status %d" % (what, status))

    def alloc(self, nbytes):
        ptr = ctypes.c_void_p()
        self._check(self.rt.hipMalloc(ctypes.byref(ptr), nbytes), "hipMalloc")
        return ptr.value

    def free(self, ptr):
        self._check(self.rt.hipFree(ptr), "hipFree")

    def upload(self, array):
        array = np.ascontiguousarray(array)
        ptr = self.alloc(array.nbytes)
        self._check(self.rt.hipMemcpy(ptr, array.ctypes.data, array.nbytes, 1), "hipMemcpy H2D")
        return ptr

    def write(self, ptr, array):
        """Copies `array` into the existing buffer at `ptr`."""
        array = np.ascontiguousarray(array)
        self._check(self.rt.hipMemcpy(ptr, array.ctypes.data, array.nbytes, 1), "hipMemcpy H2D")

    def download(self, ptr, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        self._check(self.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")
        self._check(self.rt.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2), "hipMemcpy D2H")
        return out


@pytest.fixture(scope="session")
def device_memory(polymul):
    return DeviceMemory(polymul)
