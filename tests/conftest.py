"""Test configuration: the `gpu` marker, and a build of the native parts when missing."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct MI355X (gfx950) GPU")


def _ensure_built():
    import importlib.util
    import sysconfig
    here = os.path.join(ROOT, "nwgrad_amd")
    ext = os.path.join(here, "_nwgrad_amd" + (sysconfig.get_config_var("EXT_SUFFIX") or ".so"))
    if not (os.path.exists(ext) and os.path.exists(os.path.join(here, "libnwgrad_hip.so"))):
        spec = importlib.util.spec_from_file_location("_nwg_build", os.path.join(here, "_build.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build()


_ensure_built()


@pytest.fixture(scope="session")
def gpu():
    import nwgrad_amd
    if not nwgrad_amd.device_available():
        pytest.fail("no usable gfx950 device: GPU tests must run on an MI355X")
    return True
