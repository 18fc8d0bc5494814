import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def golden():
    return os.path.join(HERE, "golden")


@pytest.fixture(scope="session")
def scenes_dir():
    return os.path.join(ROOT, "pcr_amd", "scenes")


@pytest.fixture(scope="session")
def gpu_ctx():
    import pcr_amd
    ctx = pcr_amd.Context(0)
    yield ctx
    ctx.close()
