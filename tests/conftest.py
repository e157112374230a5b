import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "newton-xpbd-vulkan_amd")
for p in (PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); runs the HIP kernels")


@pytest.fixture(scope="session")
def vk():
    import vkxpbd
    return vkxpbd


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib
