import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    import marx_amd
    lib = marx_amd.load()
    return lib.mxg_init() == 0


@pytest.fixture(scope="session")
def gpu():
    """Skips nothing: a gpu-marked test on a box without a usable GPU fails
    loudly with the library's own error message."""
    import marx_amd
    lib = marx_amd.load()
    if lib.mxg_init() != 0:
        raise RuntimeError("libmarx_gpu could not initialise the GPU: "
                           + marx_amd._lib.last_error())
    return lib
