import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "lgpsf-hessian_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")

if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
if PKG_DIR not in sys.path:
    sys.path.insert(0, PKG_DIR)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_pkg():
    """Import lgpsf-hessian_amd/ as the package `lgpsf_hessian_amd`."""
    if "lgpsf_hessian_amd" in sys.modules:
        return sys.modules["lgpsf_hessian_amd"]
    spec = importlib.util.spec_from_file_location(
        "lgpsf_hessian_amd", os.path.join(PKG_DIR, "__init__.py"),
        submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["lgpsf_hessian_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def lgh():
    return load_pkg()


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O
