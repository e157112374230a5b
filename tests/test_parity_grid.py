"""The projected ray-parity grid of the watertight mesh sign test.

The reference decides inside/outside of a watertight mesh by counting hits of
a fixed-direction ray against every triangle of the mesh
(generate_mesh_contacts.comp:372-388).  The product tests only the triangles
registered in the ray origin's cell of a grid of triangle projections along
that direction; the sign is unchanged iff every triangle the all-triangle loop
hits is registered in that cell.  This checks exactly that containment against
the oracle's all-triangle loop, on rays aimed through vertices, edges and
faces (where the watertight test is most fragile) and on random rays."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O
from vkxpbd.scenes import icosphere

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CAPI = os.path.join(ROOT, "newton-xpbd-vulkan_amd", "vkxpbd", "libvkxpbd.so")
DIR = np.array([1.029700994, 1.086472273, 1.034694910], np.float32)


def candidates(verts, tris, pts):
    L = C.CDLL(CAPI)
    fp, up, sp = C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.POINTER(C.c_size_t)
    L.xpbd_parity_grid_candidates.argtypes = [fp, C.c_size_t, up, C.c_size_t, fp, C.c_size_t, up, up, C.c_size_t, sp]
    v = np.ascontiguousarray(verts, np.float32).reshape(-1)
    t = np.ascontiguousarray(tris, np.uint32).reshape(-1)
    p = np.ascontiguousarray(pts, np.float32).reshape(-1)
    npt = p.size // 3
    counts = np.zeros(npt, np.uint32)
    cap = 256 * npt + 16
    ids = np.zeros(cap, np.uint32)
    tot = C.c_size_t(0)
    assert L.xpbd_parity_grid_candidates(v.ctypes.data_as(fp), v.size // 3, t.ctypes.data_as(up), t.size // 3,
                                         p.ctypes.data_as(fp), npt, counts.ctypes.data_as(up), ids.ctypes.data_as(up),
                                         cap, C.byref(tot)) == 0
    assert tot.value <= cap
    out, o = [], 0
    for c in counts:
        out.append(ids[o:o + c])
        o += int(c)
    return out


def probe_points(verts, tris, rng, n_each):
    """Origins whose rays pass through vertices, edge points and face points."""
    v = verts.astype(np.float64)
    t = tris.astype(np.int64)
    pick = rng.integers(0, len(v), n_each)
    through = [v[pick]]
    e = t[rng.integers(0, len(t), n_each)]
    s = rng.uniform(0, 1, (n_each, 1))
    through.append(v[e[:, 0]] * s + v[e[:, 1]] * (1 - s))
    f = t[rng.integers(0, len(t), n_each)]
    b = rng.dirichlet([1, 1, 1], n_each)
    through.append(v[f[:, 0]] * b[:, :1] + v[f[:, 1]] * b[:, 1:2] + v[f[:, 2]] * b[:, 2:])
    through = np.concatenate(through)
    back = rng.uniform(0.0, 1.5, (len(through), 1)) * np.linalg.norm(v, axis=1).max()
    pts = through - back * DIR.astype(np.float64)
    lo, hi = v.min(0), v.max(0)
    rand = rng.uniform(lo - 0.2 * (hi - lo), hi + 0.2 * (hi - lo), (n_each, 3))
    return np.concatenate([pts, rand]).astype(np.float32)


def meshes():
    v4, t4 = icosphere(4)
    out = [("ico4", v4, t4)]
    v3, t3 = icosphere(3, radius=40.0, center=(100.0, -30.0, 7.0))
    out.append(("ico3_far_large", v3, t3))
    rng = np.random.default_rng(3)
    # a thin sliver soup: triangles whose projections are nearly degenerate
    base = rng.uniform(-1, 1, size=(200, 3)).astype(np.float32)
    tri = np.stack([base, base + 1e-3 * rng.normal(size=(200, 3)), base + DIR * rng.uniform(0.1, 0.5, (200, 1))], 1)
    out.append(("slivers", tri.reshape(-1, 3).astype(np.float32), np.arange(600, dtype=np.uint32).reshape(-1, 3)))
    return out


@pytest.mark.parametrize("name,verts,tris", meshes(), ids=[m[0] for m in meshes()])
def test_parity_grid_contains_every_hit(name, verts, tris):
    rng = np.random.default_rng(11)
    pts = probe_points(np.asarray(verts, np.float32), np.asarray(tris, np.uint32), rng, 300)
    hits = O.ray_hit_lists(verts, tris, pts)
    cands = candidates(verts, tris, pts)
    n_hit = 0
    for k, (h, c) in enumerate(zip(hits, cands)):
        missing = np.setdiff1d(h, c)
        assert missing.size == 0, f"point {k}: hit triangles {missing} not registered in its cell"
        n_hit += len(h)
    assert n_hit > 100   # the probes do meet the mesh
    # the grid prunes: a cell holds a small fraction of the triangles
    mean = np.mean([len(c) for c in cands])
    assert mean < 0.25 * len(np.asarray(tris).reshape(-1, 3)) or len(np.asarray(tris).reshape(-1, 3)) < 400


def test_parity_grid_inside_outside_sphere():
    """Sign from the candidate subset equals the closed form for a sphere."""
    v, t = icosphere(3, radius=1.0, center=(0.0, 0.0, 0.0))
    rng = np.random.default_rng(5)
    d = rng.normal(size=(400, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.concatenate([rng.uniform(0.0, 0.9, 200), rng.uniform(1.05, 2.0, 200)])
    pts = (d * r[:, None]).astype(np.float32)
    hits = O.ray_hit_lists(v, t, pts)
    inside = np.array([len(h) % 2 == 1 for h in hits])
    assert np.array_equal(inside, r < 1.0)
    cands = candidates(v, t, pts)
    for h, c in zip(hits, cands):
        assert np.setdiff1d(h, c).size == 0
