import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
PKG_DIR = os.path.join(ROOT, "flowinvarianttransfer.jl_amd")


def load_package():
    """Import the backend package (its directory name is not a Python identifier)."""
    if "fit_amd" in sys.modules:
        return sys.modules["fit_amd"]
    spec = importlib.util.spec_from_file_location(
        "fit_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["fit_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def fit():
    return load_package()


@pytest.fixture(scope="session")
def oracle():
    from oracle import fit_oracle
    return fit_oracle
