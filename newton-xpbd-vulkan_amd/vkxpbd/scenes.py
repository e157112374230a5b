"""Synthetic scenes for the benchmark configurations (pure numpy).

Z is up: the reference's ground plane normal is (0, 0, 1) and hard-coded
(shaders/generate_plane_contacts.comp:37,53), so "ground plane y=0" of the
configuration text is the plane z = 0 here, with gravity (0, 0, -9.81).
Particle ids run z-major (z outer, then y, then x), the order of the hash
grid's cell ids.  Random numbers come from numpy's PCG64 with seed = config
index; lattice positions are built in float64, jittered, then cast to float32.
"""
from __future__ import annotations

import math

import numpy as np

PARTICLE_INV_MASS = 10.0   # plugin default mass 0.1 kg (newton_vulkan_plugin.py:86,384)


def lattice(nx, ny, nz, spacing, origin, jitter, rng):
    """(nx*ny*nz, 3) float32 lattice, index = (k*ny + j)*nx + i."""
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    p = np.stack([origin[0] + spacing * i, origin[1] + spacing * j, origin[2] + spacing * k], axis=-1)
    p = p.reshape(-1, 3).astype(np.float64)
    if jitter > 0:
        p += rng.uniform(-jitter, jitter, size=p.shape)
    return p.astype(np.float32)


def icosphere(level, radius=0.5, center=(0.0, 0.0, 0.5)):
    """Subdivided icosahedron: 20*4^level outward-wound triangles over
    10*4^level + 2 shared vertices (indexed, hence watertight)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(level):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        lo, hi = np.minimum(e[:, 0], e[:, 1]), np.maximum(e[:, 0], e[:, 1])
        uk, inv = np.unique(lo * nv + hi, return_inverse=True)
        mid = 0.5 * (v[uk // nv] + v[uk % nv])
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        v = np.concatenate([v, mid], axis=0)
        m = (nv + inv).reshape(3, -1)
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = m[0], m[1], m[2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)], axis=0)
    verts = (v * radius + np.asarray(center, dtype=np.float64)).astype(np.float32)
    return verts, f.astype(np.uint32)


def solver_config(vkxpbd_module, r, **overrides):
    """SolverConfig for a pile of radius-r particles: search radius 2r and
    particle_max_radius r (the plugin's rule, newton_vulkan_plugin.py:423-424),
    every other field at the reference default (src/scene.h:14-41)."""
    c = vkxpbd_module.SolverConfig()
    c.search_radius = 2.0 * r
    c.particle_max_radius = r
    c.gravity = [0.0, 0.0, -9.81]
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


def _particles(q, r, qd=None):
    n = q.shape[0]
    return dict(q=q, qd=np.zeros_like(q) if qd is None else qd, inv_mass=np.full(n, PARTICLE_INV_MASS, np.float32),
                radii=np.full(n, r, np.float32), flags=np.ones(n, np.int32))


def pile_4096(seed=0):
    """4,096 particles, 16^3 lattice, r = 0.01, spacing 2r, jitter 0.1r, plane."""
    rng = np.random.default_rng(seed)
    r = 0.01
    q = lattice(16, 16, 16, 2 * r, (-0.15, -0.15, 0.02), 0.1 * r, rng)
    return dict(name="pile_4096", r=r, particles=_particles(q, r), meshes=[], frames=100, dt=1.0 / 60.0, substeps=8,
                iterations=1)


def pile_1m_icosphere(seed=1, nx=128, ny=128, nz=64, level=6):
    """1,048,576 particles (128 x 128 x 64, r = 0.005, spacing 2r, jitter
    0.1r, bottom layer at z = 1.05) over a radius-0.5 icosphere (81,920
    triangles) centred at (0, 0, 0.5), plus the ground plane."""
    rng = np.random.default_rng(seed)
    r = 0.005
    s = 2 * r
    q = lattice(nx, ny, nz, s, (-(nx - 1) * s / 2, -(ny - 1) * s / 2, 1.05), 0.1 * r, rng)
    verts, tris = icosphere(level, 0.5, (0.0, 0.0, 0.5))
    return dict(name=f"pile_{q.shape[0]}_icosphere{tris.shape[0]}", r=r, particles=_particles(q, r),
                meshes=[(verts, tris)], frames=60, dt=1.0 / 60.0, substeps=8, iterations=1)


def pile(n, r, rng, z0=None):
    """n particles of radius r stacked on a near-cubic lattice (spacing 2r,
    jitter 0.1r) above the origin; the last layer may be partial."""
    side = max(1, int(round(n ** (1.0 / 3.0))))
    nz = -(-n // (side * side))
    q = lattice(side, side, nz, 2 * r, (-(side - 1) * r, -(side - 1) * r, r * 1.2 if z0 is None else z0), 0.1 * r, rng)
    return q[:n]


def random_quats(rng, n):
    """Uniform random unit quaternions (Shoemake), xyzw, float32."""
    u1, u2, u3 = rng.uniform(size=(3, n))
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], 1)
    return q.astype(np.float32)


def boxes_and_pile(seed=2, nb=512, n1=64):
    """nb box bodies (half-extent 0.05, density 1000: 1 kg) with uniform
    random orientations, centres uniform in [-1,1] x [-1,1] x [0.1,1.1], plus
    n1^3 particles (r = 0.005, spacing 0.01) with the bottom layer at z = 1.3."""
    rng = np.random.default_rng(seed)
    r = 0.005
    pos = np.stack([rng.uniform(-1, 1, nb), rng.uniform(-1, 1, nb), rng.uniform(0.1, 1.1, nb)], 1).astype(np.float32)
    quats = random_quats(rng, nb)
    s = 2 * r
    q = lattice(n1, n1, n1, s, (-(n1 - 1) * s / 2, -(n1 - 1) * s / 2, 1.3), 0.1 * r, rng)
    boxes = [(pos[k], quats[k], np.float32([0.05, 0.05, 0.05]), 1000.0) for k in range(nb)]
    return dict(name=f"boxes_{nb}_particles_{q.shape[0]}", r=r, particles=_particles(q, r), meshes=[], boxes=boxes,
                frames=60, dt=1.0 / 60.0, substeps=8, iterations=1)


def batch_of_piles(seed=3, num_scenes=256, lo=1024, hi=65536):
    """num_scenes independent piles, particle counts log-uniform on [lo, hi]
    (seeded), r = 0.005, each over its own ground plane and hash grid."""
    rng = np.random.default_rng(seed)
    r = 0.005
    counts = np.exp(rng.uniform(np.log(lo), np.log(hi), num_scenes)).astype(np.int64)
    parts = [pile(int(c), r, rng, z0=0.02) for c in counts]
    q = np.concatenate(parts).astype(np.float32)
    return dict(name=f"batch_{num_scenes}_scenes_{q.shape[0]}", r=r, particles=_particles(q, r), meshes=[],
                counts=counts, frames=60, dt=1.0 / 60.0, substeps=8, iterations=1)


def scene_slice(scene, k):
    """Scene k of a batched scene dict, as a plain single-scene dict."""
    off = np.concatenate([[0], np.cumsum(scene["counts"])])
    a, b = int(off[k]), int(off[k + 1])
    p = {key: v[a:b] for key, v in scene["particles"].items()}
    return dict(name=f"{scene['name']}[{k}]", r=scene["r"], particles=p, meshes=list(scene["meshes"]))


def build_batch_solver(vkxpbd_module, scene, max_contacts=0, **cfg_overrides):
    """Initialised batched XpbdSolver for a scene dict with per-scene counts (batch_of_piles)."""
    cfg = solver_config(vkxpbd_module, scene["r"], **cfg_overrides)
    s = vkxpbd_module.XpbdSolver()
    p = scene["particles"]
    s.init_batch(cfg, np.asarray(scene["counts"], np.int64), max_contacts)
    for verts, tris in scene["meshes"]:
        s.add_mesh(verts.reshape(-1), tris.reshape(-1))
    s.finalize_meshes()
    s.set_particles(p["q"], p["qd"], p["inv_mass"], p["radii"], p["flags"])
    return s


def rollout_pile_32768(seed=4):
    """32,768 particles (32^3, r = 0.01, spacing 2r, jitter 0.1r) with
    N(0, 0.5) initial velocities; 64 substeps of 1/480 s; seeded targets."""
    rng = np.random.default_rng(seed)
    r = 0.01
    q = lattice(32, 32, 32, 2 * r, (-0.31, -0.31, 0.02), 0.1 * r, rng)
    qd = rng.normal(0.0, 0.5, size=q.shape).astype(np.float32)
    target = (q.astype(np.float64) + rng.uniform(-0.05, 0.05, size=q.shape)).astype(np.float32)
    return dict(name="rollout_32768", r=r, particles=_particles(q, r, qd), meshes=[], frames=8, dt=1.0 / 60.0,
                substeps=8, iterations=1, target=target)


def build_solver(vkxpbd_module, scene, max_contacts=0, **cfg_overrides):
    """Initialised XpbdSolver for a scene dict (meshes added and finalised)."""
    cfg = solver_config(vkxpbd_module, scene["r"], **cfg_overrides)
    s = vkxpbd_module.XpbdSolver()
    p = scene["particles"]
    s.init(cfg, p["q"].shape[0], max_contacts)
    for verts, tris in scene["meshes"]:
        s.add_mesh(verts.reshape(-1), tris.reshape(-1))
    for pos, quat, half, density in scene.get("boxes", []):
        s.add_box_body(pos, quat, half, density)
    s.finalize_meshes()
    s.set_particles(p["q"], p["qd"], p["inv_mass"], p["radii"], p["flags"])
    return s
