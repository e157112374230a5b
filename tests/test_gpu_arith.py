"""The particle-particle projection takes its correctly rounded reciprocals
and square roots through shorter gfx950 sequences (xpbd_device.h rcp_cr,
sqrt_cr) wherever they equal the compiler's correctly rounded 1.0f / d and
sqrtf(a).  xpbd_check_arith compares them on every float of those ranges on
the GPU: 2^31 reciprocals, about 1.9e9 square roots and zero."""
import ctypes as C
import os

import pytest

from gpu_util import log

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "newton-xpbd-vulkan_amd", "vkxpbd", "libvkxpbd.so")


def test_short_reciprocal_and_sqrt_equal_the_correctly_rounded_ones():
    L = C.CDLL(LIB)
    L.xpbd_last_error.restype = C.c_char_p
    out = (C.c_uint64 * 3)()
    log("xpbd_check_arith: every float of the reciprocal and square-root ranges")
    rc = L.xpbd_check_arith(out)
    assert rc == 0, L.xpbd_last_error(None)
    assert list(out) == [0, 0, 0], f"mismatches (rcp, sqrt, sqrt(0)): {list(out)}"
