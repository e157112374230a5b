import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "newton-xpbd-vulkan_amd")
for p in (PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

# torch (when installed) must be loaded before libvkxpbd.so so the process has
# one HIP runtime; the package import does that, before any test loads the
# library directly through ctypes
import vkxpbd  # noqa: E402,F401


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
