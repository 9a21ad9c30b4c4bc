import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libyar_gpu.so")


@pytest.fixture(scope="session")
def scene_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("scenes")


@pytest.fixture(scope="session")
def c1_path(scene_dir):
    import yar_amd
    return yar_amd.generate_scene(1, str(scene_dir / "c1.pbrt"), seed=1)


@pytest.fixture(scope="session")
def c2_small_path(scene_dir):
    import yar_amd
    return yar_amd.generate_scene(2, str(scene_dir / "c2_small.pbrt"), seed=1, scale=0.05)


@pytest.fixture(scope="session")
def oracle():
    from tests import oracle_lib
    return oracle_lib
