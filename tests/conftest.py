import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def import_package():
    """Imports dune-pdelab_amd/ as module `dune_pdelab_amd`."""
    if "dune_pdelab_amd" in sys.modules:
        return sys.modules["dune_pdelab_amd"]
    pkg = os.path.join(ROOT, "dune-pdelab_amd")
    spec = importlib.util.spec_from_file_location("dune_pdelab_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["dune_pdelab_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def pkg():
    return import_package()


@pytest.fixture(scope="session")
def of():
    import oracle_ffi
    return oracle_ffi
