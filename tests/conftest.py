import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "path-tracer_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libptcore.so")


@pytest.fixture(scope="session")
def ctx():
    import ptcore

    c = ptcore.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def cornell():
    import ptcore

    return ptcore.World.generate(ptcore.SCENE_CORNELL, 1, 256, 256)


@pytest.fixture(scope="session")
def field():
    import ptcore

    return ptcore.World.generate(ptcore.SCENE_FIELD, 2, 1920, 1080)
