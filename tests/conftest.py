"""Shared fixtures: the product package (libagx.so through its ctypes binding) and the oracle."""
import importlib.util
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG_DIR = ROOT / "physicssimulation-agony_amd"
sys.path.insert(0, str(Path(__file__).resolve().parent))


def load_package():
    """Imports physicssimulation-agony_amd (a directory name that is not an identifier) as 'agx'."""
    if "agx" in sys.modules:
        return sys.modules["agx"]
    spec = importlib.util.spec_from_file_location("agx", PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["agx"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def agx():
    return load_package()


@pytest.fixture(scope="session")
def scenes(agx):
    import agx.scenes as s  # noqa: F401  (submodule of the loaded package)
    return sys.modules["agx.scenes"]


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib
