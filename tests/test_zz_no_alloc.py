"""xpbd_sort_pairs_noalloc, the allocation-free variant of xpbd_sort_pairs
(its temporaries live in the library's static device scratch), through the
C-ABI: two calls in a row, the second on fewer pairs, each equal bit for bit
to xpbd_sort_pairs on the same input; sizes beyond its maximum are rejected;
a call with no pairs only prepares the device."""
import ctypes as C
import os

import numpy as np
import pytest

from gpu_util import log

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "newton-xpbd-vulkan_amd", "vkxpbd", "libvkxpbd.so")
MAX_PAIRS = 1 << 20


def lib():
    L = C.CDLL(LIB)
    L.xpbd_last_error.restype = C.c_char_p
    for f in (L.xpbd_sort_pairs, L.xpbd_sort_pairs_noalloc):
        f.argtypes = [C.c_void_p] * 4 + [C.c_size_t, C.c_int, C.c_void_p]
    return L


def run(fn, ki, vi, n, key_bits):
    import torch
    ko = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    vo = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    rc = fn(ki.data_ptr(), vi.data_ptr(), ko.data_ptr(), vo.data_ptr(), n, key_bits, None)
    torch.cuda.synchronize()
    return rc, ko.cpu().numpy().view(np.uint32)[:n], vo.cpu().numpy().view(np.uint32)[:n]


def test_noalloc_sort_equals_allocating_sort():
    import torch
    L = lib()
    ki0 = torch.zeros(1, dtype=torch.int32, device="cuda")
    assert L.xpbd_sort_pairs_noalloc(ki0.data_ptr(), ki0.data_ptr(), ki0.data_ptr(), ki0.data_ptr(), 0, 24, None) == 0
    rng = np.random.default_rng(7)
    for n, key_bits in ((MAX_PAIRS, 24), (300001, 21)):
        keys = rng.integers(0, 1 << key_bits, size=n, dtype=np.uint64).astype(np.uint32)
        keys[: n // 3] = rng.integers(0, 64, size=n // 3)        # long runs of equal keys
        vals = rng.permutation(n).astype(np.uint32)
        ki = torch.from_numpy(keys.view(np.int32)).cuda()
        vi = torch.from_numpy(vals.view(np.int32)).cuda()
        log(f"noalloc sort: n={n} key_bits={key_bits}")
        rc, ka, va = run(L.xpbd_sort_pairs, ki, vi, n, key_bits)
        assert rc == 0, L.xpbd_last_error(None)
        rc, kb, vb = run(L.xpbd_sort_pairs_noalloc, ki, vi, n, key_bits)
        assert rc == 0, L.xpbd_last_error(None)
        np.testing.assert_array_equal(kb, ka)
        np.testing.assert_array_equal(vb, va)
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(vb, vals[order])


def test_noalloc_sort_rejects_sizes_beyond_its_scratch():
    import torch
    L = lib()
    t = torch.zeros(1, dtype=torch.int32, device="cuda")
    p = t.data_ptr()
    assert L.xpbd_sort_pairs_noalloc(p, p, p, p, MAX_PAIRS + 1, 24, None) == 1
    assert b"exceeds the maximum" in L.xpbd_last_error(None)
