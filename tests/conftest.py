"""Test setup: import paths for the package and the oracle, the `gpu` marker."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "cube-tracing_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libctamd.so")


@pytest.fixture(scope="session")
def gpu():
    """Fails (does not skip) when the HIP path cannot run: a GPU test must
    exercise the native library, never a fallback."""
    import ctamd
    assert os.path.exists(ctamd.LIB_PATH), "libctamd.so is not built"
    assert ctamd.device_available(), "no gfx950 device visible"
    return ctamd
