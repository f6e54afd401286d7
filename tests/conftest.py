import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def _gpu_available():
    try:
        from rocflu_amd import _native
        return _native.lib().rflu_gpu_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skips nothing: a gpu-marked test on a box without a GPU must fail loudly."""
    from rocflu_amd import _native
    n = _native.lib().rflu_gpu_device_count()
    assert n > 0, "no GPU visible to librocflu_amd.so"
    return n
