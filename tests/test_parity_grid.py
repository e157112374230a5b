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


def dist_bounds(verts, tris, cap, pts):
    L = C.CDLL(CAPI)
    fp, up = C.POINTER(C.c_float), C.POINTER(C.c_uint32)
    L.xpbd_dist_grid_bounds.argtypes = [fp, C.c_size_t, up, C.c_size_t, C.c_float, fp, C.c_size_t, fp]
    v = np.ascontiguousarray(verts, np.float32).reshape(-1)
    t = np.ascontiguousarray(tris, np.uint32).reshape(-1)
    p = np.ascontiguousarray(pts, np.float32).reshape(-1)
    out = np.zeros(p.size // 3, np.float32)
    assert L.xpbd_dist_grid_bounds(v.ctypes.data_as(fp), v.size // 3, t.ctypes.data_as(up), t.size // 3, cap,
                                   p.ctypes.data_as(fp), p.size // 3, out.ctypes.data_as(fp)) == 0
    return out


def exact_min_dist(verts, tris, pts):
    """Brute-force point-to-triangle distance (double, textbook Voronoi-region
    closest point), minimum over all triangles."""
    v = np.asarray(verts, np.float64)
    t = np.asarray(tris, np.int64).reshape(-1, 3)
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    out = []
    for p in np.asarray(pts, np.float64):
        best = np.inf
        for k in range(len(t)):
            cp = _closest(p, a[k], b[k], c[k])
            best = min(best, float(np.linalg.norm(p - cp)))
        out.append(best)
    return np.array(out)


def _closest(p, a, b, c):
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + ab * (d1 / (d1 - d3))
    cp_ = p - c
    d5, d6 = ab @ cp_, ac @ cp_
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + ac * (d2 / (d2 - d6))
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)))
    den = 1.0 / (va + vb + vc)
    return a + ab * (vb * den) + ac * (vc * den)


def test_distance_bounds_are_lower_bounds():
    """The mesh query skips a particle when its cell's bound exceeds the search
    distance: every bound must be <= the exact distance to the mesh, and a
    point outside the grid must be farther than the cap."""
    v, t = icosphere(2, radius=0.5, center=(0.0, 0.0, 0.5))
    cap = 0.05
    rng = np.random.default_rng(8)
    d = rng.normal(size=(600, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(0.3, 0.7, 600)
    pts = np.concatenate([d * r[:, None] + [0, 0, 0.5], rng.uniform(-0.7, 0.7, (200, 3)) + [0, 0, 0.5]])
    pts = pts.astype(np.float32)
    lb = dist_bounds(v, t, cap, pts)
    ex = exact_min_dist(v, t, pts)
    inside = lb >= 0
    assert np.all(lb[inside] <= ex[inside] + 1e-7)
    assert np.all(ex[~inside] > cap)
    # the bound is useful: tight to about a cell diagonal near the surface
    near = inside & (ex < cap)
    assert np.all(ex[near] - lb[near] <= 0.05 * np.sqrt(3) * 0.5 + 0.06)
    assert np.count_nonzero(inside & (lb > 0.01)) > 100
