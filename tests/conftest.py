import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_pkg():
    """Import the block-dda_amd package (its directory name has a hyphen)."""
    name = "block_dda_amd"
    if name in sys.modules:
        return sys.modules[name]
    path = os.path.join(ROOT, "block-dda_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location(name, path,
                                                  submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def bdda():
    return load_pkg()


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
