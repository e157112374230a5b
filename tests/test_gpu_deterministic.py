"""Deterministic mode: every sum the reference takes with float atomics (a
particle's shape deltas and restitution updates, a body's deltas) is taken in
the order of a sequential pass over the contacts as the generators emit them,
which is the oracle's order.  Where the atomic order could change bits -- rigid
boxes with many particle and corner contacts, restitution on particles with
two shape contacts -- the deterministic solver equals the oracle bit for bit;
where it could not, it equals the default mode."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import assert_contacts_equal, gpu_contacts, log, make_pair, scene_from
from vkxpbd import scenes

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def exact(g, o, what, bodies=False):
    gp, op = g.get_positions().astype(np.float64), o.get_positions()
    gv, ov = g.get_velocities().astype(np.float64), o.get_velocities()
    msg = f"  {what}: max|dx|={float(np.max(np.abs(gp - op))):.3e} max|dv|={float(np.max(np.abs(gv - ov))):.3e}"
    if bodies:
        db = float(np.max(np.abs(g.get_body_q().astype(np.float64) - o.get_body_q())))
        msg += f" max|dq_body|={db:.3e}"
    log(msg)
    np.testing.assert_array_equal(gp, op, err_msg=what)
    np.testing.assert_array_equal(gv, ov, err_msg=what)
    if bodies:
        np.testing.assert_array_equal(g.get_body_q().astype(np.float64), o.get_body_q(), err_msg=what)
        np.testing.assert_array_equal(g.get_body_qd().astype(np.float64), o.get_body_qd(), err_msg=what)


def box_scene(nb_side=3, seed=12):
    rng = np.random.default_rng(seed)
    r = 0.005
    nb = nb_side ** 3
    pos = np.stack(np.meshgrid(*(np.arange(nb_side),) * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = (pos * [0.25, 0.25, 0.2] + [-0.25, -0.25, 0.1]).astype(np.float32)
    quats = scenes.random_quats(rng, nb)
    q = scenes.lattice(24, 24, 6, 2 * r, (-0.115, -0.115, 0.7), 0.1 * r, rng)
    sc = scene_from(q, r)
    sc["boxes"] = [(pos[k], quats[k], np.float32([0.05, 0.05, 0.05]), 1000.0) for k in range(nb)]
    return sc


def test_boxes_and_particles_equal_oracle(vk):
    """27 tumbling boxes with particles raining onto them and their corners
    hitting the plane: bit-exact against the oracle for 40 frames."""
    sc = box_scene()
    g, o = make_pair(vk, sc)
    g.set_deterministic(True)
    assert g.deterministic
    for f in range(1, 41):
        g.step(DT, 8, 1)
        o.step(DT, 8, 1)
        assert g.last_contact_count == o.last_contact_count
        if f in (1, 2, 5, 10, 20, 40):
            exact(g, o, f"27 boxes frame {f}", bodies=True)


def test_512_boxes_and_262144_particles_equal_oracle(vk):
    """The rigid-body workload at full size (512 boxes, 262,144 particles),
    which the default mode matches only to 1e-5 m: bit-exact in
    deterministic mode, frame after frame."""
    sc = scenes.boxes_and_pile()
    g, o = make_pair(vk, sc)
    g.set_deterministic(True)
    for f in range(1, 9):
        g.step(DT, 8, 1)
        o.step(DT, 8, 1)
        assert g.last_contact_count == o.last_contact_count
        exact(g, o, f"512 boxes + pile frame {f}", bodies=True)


def test_restitution_with_two_shape_contacts_equals_oracle(vk):
    """Particles in the wedge between a sphere and the ground touch both, with
    restitution on: each particle's two restitution updates read the velocity
    the previous one left (the reference's atomics on the velocity they read),
    in shape order; bit-exact against the oracle."""
    rng = np.random.default_rng(31)
    r = 0.005
    n = 2000
    ang = rng.uniform(0.0, 2.0 * np.pi, n)
    z = rng.uniform(0.004, 0.012, n)
    rho = np.sqrt(0.5 ** 2 - (0.5 - z) ** 2) + rng.uniform(0.0, 0.012, n)
    q = np.stack([rho * np.cos(ang), rho * np.sin(ang), z], 1).astype(np.float32)
    qd = rng.normal(0.0, 0.3, size=q.shape).astype(np.float32)
    qd[:, 2] -= 0.5
    verts, tris = scenes.icosphere(4)
    sc = scene_from(q, r, qd, meshes=[(verts, tris)])
    g, o = make_pair(vk, sc, enable_restitution=True, soft_contact_restitution=0.6)
    g.set_deterministic(True)
    for f in range(1, 5):
        g.step(DT, 4, 1)
        o.step(DT, 4, 1)
        assert g.last_contact_count == o.last_contact_count
        exact(g, o, f"wedge with restitution frame {f}")
    c = gpu_contacts(g)
    both = set(c["particles"][c["shapes"] == 0].tolist()) & set(c["particles"][c["shapes"] == 1].tolist())
    log(f"  particles with a plane and a mesh contact: {len(both)}")
    assert len(both) > 50
    assert_contacts_equal(c, o.get_contacts())


def test_deterministic_equals_default_where_order_is_free(vk):
    """The 4,096-particle pile over a small sphere: at most two shape contacts
    per particle and no bodies, so the default mode's atomics are exact too;
    both modes give the same bits, eager and replayed."""
    sc = scenes.pile_4096()
    v, t = scenes.icosphere(2, 0.1, (0.0, 0.0, 0.1))
    sc["meshes"] = [(v, t)]
    a = scenes.build_solver(vk, sc)
    b = scenes.build_solver(vk, sc)
    b.set_deterministic(True)
    for f in range(20):
        a.step(DT, 8, 1)
        b.step(DT, 8, 1)
        np.testing.assert_array_equal(a.get_positions(), b.get_positions(), err_msg=f"frame {f}")
        np.testing.assert_array_equal(a.get_velocities(), b.get_velocities(), err_msg=f"frame {f}")
    log(f"deterministic == default over 20 frames, contacts {a.last_contact_count}")
