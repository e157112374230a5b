"""xpbd_sort_pairs (the stand-alone stable radix sort, GpuSort's counterpart,
src/gpu_sort.h:18-129) through the C-ABI on device buffers: the output equals
numpy's stable argsort by key (equal keys keep their input order), the inputs
are untouched, across tile-size boundaries and digit shapes."""
import ctypes as C
import os

import numpy as np
import pytest

from gpu_util import log

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "newton-xpbd-vulkan_amd", "vkxpbd", "libvkxpbd.so")


def lib():
    L = C.CDLL(LIB)
    L.xpbd_last_error.restype = C.c_char_p
    L.xpbd_sort_pairs.argtypes = [C.c_void_p] * 4 + [C.c_size_t, C.c_int, C.c_void_p]
    return L


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).cuda()


def host(t):
    return t.cpu().numpy().view(np.uint32)


def sort_on_gpu(fn, keys, vals, key_bits):
    import torch
    n = keys.size
    ki, vi = dev(keys), dev(vals)
    ko = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    vo = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    rc = fn(ki.data_ptr(), vi.data_ptr(), ko.data_ptr(), vo.data_ptr(), n, key_bits, None)
    torch.cuda.synchronize()
    return rc, host(ki), host(vi), host(ko)[:n], host(vo)[:n]


@pytest.mark.parametrize("n,key_bits,dense", [(1, 24, False), (1000, 24, True), (262143, 21, False),
                                               (262144, 24, True), (1048583, 24, False), (300000, 32, False),
                                               (5000, 7, True), (70000, 27, False)])
def test_sort_pairs_matches_stable_argsort(n, key_bits, dense):
    L = lib()
    rng = np.random.default_rng(n + key_bits)
    hi = (1 << key_bits) if not dense else min(1 << key_bits, max(2, n // 8))
    keys = rng.integers(0, hi, size=n, dtype=np.uint64).astype(np.uint32)
    vals = rng.permutation(n).astype(np.uint32)
    log(f"sort_pairs n={n} key_bits={key_bits} dense={dense}")
    rc, ki, vi, ko, vo = sort_on_gpu(L.xpbd_sort_pairs, keys, vals, key_bits)
    assert rc == 0, L.xpbd_last_error(None)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(ko, keys[order])
    np.testing.assert_array_equal(vo, vals[order])
    np.testing.assert_array_equal(ki, keys)
    np.testing.assert_array_equal(vi, vals)
