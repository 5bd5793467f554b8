import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def load_ppl():
    """Load pelican-lofar_amd/ as the package pelican_lofar_amd."""
    name = "pelican_lofar_amd"
    if name in sys.modules:
        return sys.modules[name]
    pkg_dir = os.path.join(ROOT, "pelican-lofar_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(pkg_dir, "__init__.py"),
                                                  submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def orc():
    import oracle_py
    oracle_py.build()
    return oracle_py


@pytest.fixture(scope="session")
def ppl():
    import torch
    mod = load_ppl()
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU")
    mod.init(0)
    return mod
