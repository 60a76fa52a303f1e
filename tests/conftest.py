import importlib.util
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "dune-copasi_amd"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP path)")


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path, submodule_search_locations=[str(path.parent)]
                                                  if path.name == "__init__.py" else None)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def dca():
    """The product package (ctypes binding of libdcopasi.so)."""
    if "dune_copasi_amd" in sys.modules:
        return sys.modules["dune_copasi_amd"]
    return _load("dune_copasi_amd", PKG / "__init__.py")


@pytest.fixture(scope="session")
def wl():
    return _load("dca_workloads", PKG / "workloads.py")


@pytest.fixture(scope="session")
def oracle():
    from tests import oracle_lib
    oracle_lib.load()
    return oracle_lib
