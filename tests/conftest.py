import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "blackhole-sim_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libbhgrrt.so")


@pytest.fixture(scope="session")
def renderer():
    from bhgrrt.render import Renderer

    r = Renderer(int(os.environ.get("BH_DEVICE", "0")))
    yield r
    r.close()
