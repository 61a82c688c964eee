import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "model-based-seismic-data-redatuming-with-daws_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)


def load_pkg():
    """Import the package directory (its name is not a Python identifier) as daw_amd."""
    if "daw_amd" in sys.modules:
        return sys.modules["daw_amd"]
    spec = importlib.util.spec_from_file_location("daw_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["daw_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def daw():
    return load_pkg()


@pytest.fixture(scope="session")
def gpu_daw():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("a gpu-marked test ran without a visible GPU")
    return load_pkg()
