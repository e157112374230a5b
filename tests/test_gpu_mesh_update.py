"""update_mesh on the device (animated colliders, scene.cpp:617-696): the BVH
is refitted and the narrowphase's distance-bound and parity grids rebuilt in
place.  A moved mesh must give exactly the oracle's contacts after the
oracle's own update_mesh (which refits, keeping the topology: a solver
finalized from the moved vertices builds a different tree, whose visiting
order can resolve ties between equidistant faces differently), its BVH boxes
must equal the oracle's refit, captured frames must stay valid, and a mesh
that outgrows its grids must fall back to a host rebuild with the same
results."""
import time

import numpy as np
import pytest

import oracle_lib as O
from gpu_util import assert_contacts_equal, gpu_contacts, log, make_pair, scene_from
from vkxpbd import scenes

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
CENTER = np.float32([0.0, 0.0, 0.5])


def moved(verts, angle, shift, scale=1.0):
    """The mesh rotated about z through its centre, scaled and shifted."""
    c, s = np.cos(angle), np.sin(angle)
    R = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
    v = (verts.astype(np.float64) - CENTER) @ R.T * scale + CENTER + np.asarray(shift)
    return v.astype(np.float32)


def shell(rng, n, center, radius, width=0.02):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return (d * (radius + rng.uniform(-width, width, size=(n, 1))) + center).astype(np.float32)


@pytest.mark.parametrize("level", [3, 6])
def test_moved_mesh_equals_fresh_build(vk, level):
    """A rotated and shifted icosphere (1,280 and 81,920 triangles): after
    update_mesh, the BVH boxes and collide_only's contacts equal the oracle's
    after its update_mesh, bit for bit.  Also logs the median update_mesh
    time."""
    verts, tris = scenes.icosphere(level)
    rng = np.random.default_rng(5 + level)
    r = 0.005
    shift = np.float32([0.03, -0.02, 0.015])
    v1 = moved(verts, 0.3, shift)
    q = shell(rng, 40000, CENTER + shift, 0.5)
    g, o = make_pair(vk, scene_from(q, r, meshes=[(verts, tris)]))
    g.collide_only()
    times = []
    for k in range(6):   # back and forth, ending on v1
        v = v1 if k % 2 == 1 else moved(verts, 0.29, shift * 0.9)
        t0 = time.perf_counter()
        g.update_mesh(0, v.reshape(-1))
        times.append(time.perf_counter() - t0)
    log(f"icosphere({level}), {tris.shape[0]} triangles: update_mesh median {1e3 * np.median(times):.2f} ms")
    o.update_mesh(0, v1)
    lo_g, hi_g = g.get_mesh_bvh_lowers(0), g.get_mesh_bvh_uppers(0)
    lo_o, hi_o = o.get_mesh_bvh(0)[:2]
    np.testing.assert_array_equal(lo_g.view(np.uint32), lo_o.view(np.uint32))
    np.testing.assert_array_equal(hi_g.view(np.uint32), hi_o.view(np.uint32))
    g.collide_only()
    t0 = time.perf_counter()
    o.collide_only()
    log(f"  contacts {g.last_contact_count} (oracle {time.perf_counter() - t0:.1f} s)")
    assert g.last_contact_count == o.last_contact_count > 10000
    assert_contacts_equal(gpu_contacts(g), o.get_contacts())


def test_mesh_outgrowing_its_grids_falls_back_to_a_host_rebuild(vk):
    """Scaled threefold, the icosphere no longer fits its grids' room: the
    update rebuilds on the host; a later small move goes back to the device
    path.  Both equal the oracle bit for bit."""
    verts, tris = scenes.icosphere(4)
    rng = np.random.default_rng(8)
    r = 0.005
    q = shell(rng, 30000, CENTER + np.float32([0.0, 0.0, 0.2]), 1.5, width=0.03)
    g, o = make_pair(vk, scene_from(q, r, meshes=[(verts, tris)]))
    for k, v in enumerate((moved(verts, 0.1, (0.0, 0.0, 0.2), scale=3.0),
                           moved(verts, 0.12, (0.01, 0.0, 0.2), scale=3.0))):
        g.update_mesh(0, v.reshape(-1))
        o.update_mesh(0, v)
        g.collide_only()
        o.collide_only()
        assert g.last_contact_count == o.last_contact_count > 10000
        assert_contacts_equal(gpu_contacts(g), o.get_contacts())
        np.testing.assert_array_equal(g.get_mesh_bvh_lowers(0).view(np.uint32), o.get_mesh_bvh(0)[0].view(np.uint32))


def test_frames_with_a_moving_collider_replay_and_match_the_oracle(vk):
    """A rising, turning icosphere under a falling lattice: update_mesh
    before every frame keeps the captured frames (graph replay equals eager
    launches) and every frame equals the oracle bit for bit."""
    rng = np.random.default_rng(9)
    r = 0.01
    q = scenes.lattice(16, 16, 4, 2 * r, (-0.15, -0.15, 1.02), 0.1 * r, rng)
    verts, tris = scenes.icosphere(3)
    sc = scene_from(q, r, meshes=[(verts, tris)])
    g, o = make_pair(vk, sc)
    e = scenes.build_solver(vk, sc)
    e.set_graph_mode(False)
    vel = np.tile(np.float32([0.0, 0.0, 0.6]), (verts.shape[0], 1))
    for f in range(12):
        v = moved(verts, 0.02 * f, (0.0, 0.0, 0.01 * f))
        for s in (g, e):
            s.update_mesh(0, v.reshape(-1), vel.reshape(-1))
            s.step(DT, 8, 1)
        o.update_mesh(0, v, vel)
        o.step(DT, 8, 1)
        np.testing.assert_array_equal(g.get_positions(), e.get_positions(), err_msg=f"frame {f}")
        np.testing.assert_array_equal(g.get_positions().astype(np.float64), o.get_positions(), err_msg=f"frame {f}")
        np.testing.assert_array_equal(g.get_velocities().astype(np.float64), o.get_velocities(), err_msg=f"frame {f}")
        assert g.last_contact_count == o.last_contact_count
    assert g.last_contact_count > 0
    assert_contacts_equal(gpu_contacts(g), o.get_contacts())
