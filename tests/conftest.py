import importlib.util
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "tests"))
sys.path.insert(0, os.path.join(REPO, "tests", "golden"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_dkt():
    """Imports the package from its directory (its name is not a Python identifier)."""
    if "dkt" in sys.modules:
        return sys.modules["dkt"]
    spec = importlib.util.spec_from_file_location("dkt", os.path.join(REPO, "dendro-kt_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["dkt"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def dkt():
    m = load_dkt()
    m.load()
    return m
