import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X (gfx950) GPU")


def _built():
    """Load what is built; build only when nothing is."""
    from gcorr_amd import _lib
    from oracle import fxref
    if not os.path.exists(_lib.LIB_PATH) or not os.path.exists(fxref.LIB_PATH):
        from gcorr_amd import build
        build.build_all()


_built()


@pytest.fixture(scope="session")
def gpu():
    """The GPU tests need the gfx950 device; fail loudly if the library cannot use it."""
    import gcorr_amd
    from gcorr_amd import fxkernel
    k = fxkernel.FxKernel(2, 64, 4, 2, 1.65e9, 16e6)  # raises GxcError if no usable device
    k.close()
    return True
