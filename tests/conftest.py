import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "cdbg-builder_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def _load_pkg():
    path = os.path.join(PKG_DIR, "cdbg_gpu.py")
    spec = importlib.util.spec_from_file_location("cdbg_gpu", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cdbg_gpu"] = mod
    spec.loader.exec_module(mod)
    return mod


cdbg_gpu = _load_pkg()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def gpu_ctx():
    ctx = cdbg_gpu.Context(0)
    yield ctx
    ctx.close()
