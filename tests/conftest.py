import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
PKG_DIR = os.path.join(ROOT, "nwchem-ccsd-trpdrv_amd")


def load_package():
    """Imports nwchem-ccsd-trpdrv_amd/ (a directory name with dashes) as a module."""
    name = "nwchem_ccsd_trpdrv_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def cc():
    return load_package()


@pytest.fixture(autouse=True)
def fresh_tuning(cc):
    """Every test tunes afresh: the autotuner keeps its picks per list and shape
    for the life of the process (ccsdt_tune_cache_*), so a test that tunes a list
    an earlier test tuned would otherwise get the kept pick and no timings."""
    cc.tune_cache_clear()
    yield


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.olib()
    return oracle
