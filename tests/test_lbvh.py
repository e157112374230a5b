"""LBVH (mesh_lbvh.h / mesh_bvh.h) parity: the oracle's restatement and the
product's host builder against the reference's own code compiled from its
sources (oracle/_ref), bit for bit, plus tree invariants."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O
from vkxpbd.scenes import icosphere

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CAPI = os.path.join(ROOT, "newton-xpbd-vulkan_amd", "vkxpbd", "libvkxpbd.so")


def product_lbvh(verts, tris):
    L = C.CDLL(CAPI)
    fp, up = C.POINTER(C.c_float), C.POINTER(C.c_uint32)
    L.xpbd_lbvh_build.argtypes = [fp, C.c_size_t, up, C.c_size_t, fp, fp, up, up]
    v = np.ascontiguousarray(verts, np.float32).reshape(-1)
    t = np.ascontiguousarray(tris, np.uint32).reshape(-1)
    nt = t.size // 3
    lo = np.zeros((2 * nt, 4), np.float32)
    hi = np.zeros((2 * nt, 4), np.float32)
    prim = np.zeros(nt, np.uint32)
    root = np.zeros(1, np.uint32)
    assert L.xpbd_lbvh_build(v.ctypes.data_as(fp), v.size // 3, t.ctypes.data_as(up), nt, lo.ctypes.data_as(fp),
                             hi.ctypes.data_as(fp), prim.ctypes.data_as(up), root.ctypes.data_as(up)) == 0
    return lo, hi, prim, int(root[0])


def product_refit(verts, new_verts, tris):
    L = C.CDLL(CAPI)
    fp, up = C.POINTER(C.c_float), C.POINTER(C.c_uint32)
    L.xpbd_lbvh_refit.argtypes = [fp, fp, C.c_size_t, up, C.c_size_t, fp, fp]
    v = np.ascontiguousarray(verts, np.float32).reshape(-1)
    w = np.ascontiguousarray(new_verts, np.float32).reshape(-1)
    t = np.ascontiguousarray(tris, np.uint32).reshape(-1)
    nt = t.size // 3
    lo = np.zeros((2 * nt, 4), np.float32)
    hi = np.zeros((2 * nt, 4), np.float32)
    assert L.xpbd_lbvh_refit(v.ctypes.data_as(fp), w.ctypes.data_as(fp), v.size // 3, t.ctypes.data_as(up), nt,
                             lo.ctypes.data_as(fp), hi.ctypes.data_as(fp)) == 0
    return lo, hi


def meshes():
    out = [("ico2", *icosphere(2)), ("ico4", *icosphere(4))]
    rng = np.random.default_rng(7)
    # triangle soup with many identical centroids: exercises equal Morton codes
    # and the parity tie-break of the parent choice (mesh_lbvh.h:163-166)
    base = rng.uniform(-1, 1, size=(40, 3, 3)).astype(np.float32)
    soup = np.concatenate([base, base, base[:13]], axis=0)
    out.append(("soup_dupes", soup.reshape(-1, 3), np.arange(soup.shape[0] * 3, dtype=np.uint32).reshape(-1, 3)))
    # a long thin strip: deep trees (leaf collapse at depth >= 32 is possible)
    n = 3000
    x = np.cumsum(rng.uniform(0.0, 1e-3, size=n)) ** 3
    v = np.stack([x, np.zeros(n), np.zeros(n)], 1)
    v = np.concatenate([v, v + [0, 1e-4, 0], v + [0, 0, 1e-4]]).astype(np.float32)
    t = np.stack([np.arange(n), np.arange(n) + n, np.arange(n) + 2 * n], 1).astype(np.uint32)
    out.append(("strip", v, t))
    out.append(("single", np.eye(3, dtype=np.float32), np.array([[0, 1, 2]], np.uint32)))
    return out


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


needs_ref = pytest.mark.skipif(O.ref_lib() is None, reason="oracle/_ref not built (reference sources absent)")


@needs_ref
@pytest.mark.parametrize("name,verts,tris", meshes(), ids=[m[0] for m in meshes()])
def test_oracle_lbvh_matches_reference_code(name, verts, tris):
    ref = O.ref_lbvh_build(verts, tris)
    orc = O.lbvh_build(verts, tris)
    for a, b in zip(ref[:3], orc[:3]):
        np.testing.assert_array_equal(bits(a), bits(b))
    assert ref[3] == orc[3]


@needs_ref
@pytest.mark.parametrize("name,verts,tris", meshes(), ids=[m[0] for m in meshes()])
def test_product_lbvh_matches_reference_code(name, verts, tris):
    ref = O.ref_lbvh_build(verts, tris)
    prod = product_lbvh(verts, tris)
    for a, b in zip(ref[:3], prod[:3]):
        np.testing.assert_array_equal(bits(a), bits(b))
    assert ref[3] == prod[3]


@needs_ref
def test_refit_matches_reference_code():
    v, t = icosphere(3)
    rng = np.random.default_rng(3)
    w = (v * 1.3 + rng.normal(0, 0.01, v.shape) + [0.2, -0.1, 0.4]).astype(np.float32)
    ref = O.ref_lbvh_refit(v, w, t)
    prod = product_refit(v, w, t)
    for a, b in zip(ref, prod):
        np.testing.assert_array_equal(bits(a), bits(b))


def reachable(lo, hi, root):
    out, st = [], [root]
    while st:
        k = st.pop()
        out.append(k)
        lw = int(bits(lo[k])[3])
        if not (lw >> 30) & 3:
            st += [lw & 0x3FFFFFFF, int(bits(hi[k])[3])]
    return out


@pytest.mark.parametrize("name,verts,tris", meshes(), ids=[m[0] for m in meshes()])
def test_lbvh_invariants(name, verts, tris):
    """Every primitive in exactly one reachable leaf; child boxes inside the
    parent; leaf boxes cover their triangles; leaves hold <= 4 primitives
    unless at depth >= 32."""
    lo, hi, prim, root = product_lbvh(verts, tris)
    nt = tris.shape[0]
    assert sorted(prim.tolist()) == list(range(nt))
    covered = np.zeros(nt, int)
    st = [(root, 1)]
    while st:
        k, d = st.pop()
        lw, uw = int(bits(lo[k])[3]), int(bits(hi[k])[3])
        if (lw >> 30) & 3:
            a, b = lw & 0x3FFFFFFF, uw & 0x3FFFFFFF
            assert b - a <= 4 or d >= 32
            for pi in range(a, b):
                covered[pi] += 1
                tv = verts[tris[prim[pi]]]
                assert np.all(tv.min(0) >= lo[k][:3]) and np.all(tv.max(0) <= hi[k][:3])
        else:
            for c in (lw & 0x3FFFFFFF, uw):
                assert np.all(lo[c][:3] >= lo[k][:3]) and np.all(hi[c][:3] <= hi[k][:3])
                st.append((c, d + 1))
    assert np.all(covered == 1)
