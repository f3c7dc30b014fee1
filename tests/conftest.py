import importlib.util
import sys
from pathlib import Path

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
