import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers", "slow: long-running CPU check")
    # the engine library is built in-tree; build it only when nothing is built
    # and a compiler is present (a checkout that has never run the build)
    from ultraballpit_amd import build
    if not os.path.exists(build.LIB):
        try:
            build.hipcc()
        except RuntimeError:
            return
        build.build_all()


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.lib()
    return pyoracle
