import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def _ensure_built():
    """Tests load what is built and build only when nothing is."""
    from openclrenderer_amd import build as b
    if not os.path.exists(os.path.join(ROOT, "oracle", "_build", "libocr_oracle.so")):
        b.build_oracle()
    if not os.path.exists(b.LIB):
        b.build()


@pytest.fixture(scope="session")
def oracle():
    _ensure_built()
    from oracle.binding import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def gpu_ctx():
    _ensure_built()
    from openclrenderer_amd import runtime
    ctx = runtime.Context(0)
    yield ctx
    ctx.close()
