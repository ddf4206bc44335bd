import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


def load_package():
    """Import fortkernels-_amd (a hyphenated directory) as module `fortkernels_amd`."""
    if "fortkernels_amd" in sys.modules:
        return sys.modules["fortkernels_amd"]
    pkg_dir = os.path.join(ROOT, "fortkernels-_amd")
    spec = importlib.util.spec_from_file_location(
        "fortkernels_amd", os.path.join(pkg_dir, "__init__.py"), submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["fortkernels_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def fk():
    return load_package()


@pytest.fixture(scope="session")
def fko():
    from oracle import fko as _fko
    return _fko
