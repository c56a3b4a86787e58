import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (os.path.join(ROOT, "p-trans_amd"), ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libptrans.so")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One libptrans context for the GPU tests; fails loudly if the library or the GPU
    is missing (there is no CPU fallback)."""
    from ptrans.trace import Context
    ctx = Context(seed=0, device=0)
    yield ctx
    ctx.close()
