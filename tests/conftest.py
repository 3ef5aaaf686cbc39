import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "tfg-pathtracer_amd"))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU check")


@pytest.fixture(scope="session")
def has_gpu():
    import eleven
    return eleven.device_count() > 0


@pytest.fixture(scope="session")
def gpu_ctx():
    import eleven
    if eleven.device_count() == 0:
        pytest.fail("no HIP device visible: the tests marked gpu need an MI355X")
    ctx = eleven.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def scene_root(tmp_path_factory):
    return tmp_path_factory.mktemp("scenes")
