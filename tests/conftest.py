import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "collision-detection_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def load_package():
    name = "collision_detection_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through the C ABI)")


@pytest.fixture(scope="session")
def cdx():
    return load_package()


@pytest.fixture(scope="session")
def scenes():
    load_package()
    import collision_detection_amd.scenes as s
    return s


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def ctx(cdx):
    if not os.path.exists(cdx.LIB_PATH):
        pytest.fail("libcdx.so is not built (run __graft_entry__.build())")
    if cdx.device_count() < 1:
        pytest.fail("no GPU visible: GPU tests must run on an MI355X")
    return cdx.Context(0)
