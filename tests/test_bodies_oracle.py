"""Pins for the rigid-body restatement (integrate_bodies.comp,
solve_body_contact_positions.comp, apply_body_deltas.comp and the
builder-written box contact generators) on the CPU: closed forms and
conservation laws."""
import numpy as np

from test_oracle_pins import cfg
import oracle_lib as O


def box_solver(bodies, particles=None, fp64=True, r=0.005, **kw):
    o = O.OracleSolver(fp64=fp64)
    q = np.zeros((1, 3), np.float32) + [50.0, 50.0, 50.0] if particles is None else np.asarray(particles, np.float32)
    o.init(cfg(search_radius=2 * r, particle_max_radius=r, **kw), q.shape[0])
    for b in bodies:
        o.add_box_body(*b)
    o.finalize_meshes()
    n = q.shape[0]
    o.set_particles(q, None, np.full(n, 10.0, np.float32), np.full(n, r, np.float32), np.ones(n, np.int32))
    return o


def test_rigid_free_fall_closed_form():
    """No contacts: symplectic Euler gives x_k = x0 + g h^2 k(k+1)/2, v_k = k g h;
    orientation is untouched when w = 0."""
    rot = np.array([0.2, -0.3, 0.1, 0.9]); rot /= np.linalg.norm(rot)
    o = box_solver([([0.0, 0.0, 5.0], rot, [0.05, 0.05, 0.05], 1000.0)])
    o.set_body_params(angular_damping=0.0)
    o.step(1.0 / 60.0, 8, 1)
    h = np.float64(np.float32(np.float32(1.0 / 60.0) / 8))
    q, qd = o.get_body_q()[0], o.get_body_qd()[0]
    np.testing.assert_allclose(q[:3], [0, 0, 5.0 - 9.81 * h * h * 36], atol=1e-9)
    np.testing.assert_allclose(qd[:3], [0, 0, -9.81 * 8 * h], rtol=1e-6)
    np.testing.assert_allclose(q[3:], np.float32(rot), atol=1e-7)


def test_torque_free_spin_of_a_cube():
    """A cube's inertia is isotropic, so w x Iw = 0: with no damping the spin
    is constant and the quaternion stays unit length; the body turns by |w| t."""
    o = box_solver([([0.0, 0.0, 5.0], [0, 0, 0, 1], [0.05, 0.05, 0.05], 1000.0)], gravity=[0, 0, 0])
    o.set_body_params(angular_damping=0.0)
    w = np.array([0.0, 0.0, 2.0])
    o.set_body_velocities(np.array([[0, 0, 0, *w]]))
    for _ in range(30):
        o.step(1.0 / 60.0, 8, 1)
    q, qd = o.get_body_q()[0], o.get_body_qd()[0]
    np.testing.assert_allclose(np.linalg.norm(q[3:]), 1.0, atol=1e-12)
    np.testing.assert_allclose(qd[3:], w, atol=1e-12)
    ang = 2 * np.arctan2(q[5], q[6])
    np.testing.assert_allclose(ang, 2.0 * 0.5, rtol=2e-3)   # first-order quaternion update


def test_box_settles_on_plane():
    """A box dropped flat from 2 cm settles on the plane: centre height -> the
    half extent (soft XPBD contacts leave millimetre penetration), velocity ~0."""
    o = box_solver([([0.0, 0.0, 0.07], [0, 0, 0, 1], [0.05, 0.05, 0.05], 1000.0)])
    for _ in range(120):
        o.step(1.0 / 60.0, 8, 1)
    q, qd = o.get_body_q()[0], o.get_body_qd()[0]
    assert abs(q[2] - 0.05) < 3e-3, q
    assert np.linalg.norm(qd) < 0.05, qd
    np.testing.assert_allclose(q[3:], [0, 0, 0, 1], atol=1e-6)   # symmetric contact: no tilt


def test_tilted_box_rights_itself():
    """A box landing on an edge tips onto a face: the final orientation is a
    face-down orientation (one body axis parallel to z)."""
    a = 0.35
    o = box_solver([([0.0, 0.0, 0.12], [np.sin(a / 2), 0, 0, np.cos(a / 2)], [0.05, 0.05, 0.05], 1000.0)],
                   fp64=True)
    for _ in range(180):
        o.step(1.0 / 60.0, 8, 1)
    q = o.get_body_q()[0]
    x, y, z, w = q[3:]
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    assert np.max(np.abs(R[2])) > 0.999, R
    assert abs(q[2] - 0.05) < 4e-3


def test_particle_rests_on_box():
    """A particle dropped on a static-mass box comes to rest on its top face."""
    o = box_solver([([0.0, 0.0, 0.05], [0, 0, 0, 1], [0.05, 0.05, 0.05], 1000.0, 2)],   # kinematic box
                   particles=[[0.01, -0.02, 0.2]], r=0.005)
    for _ in range(90):
        o.step(1.0 / 60.0, 8, 1)
    p = o.get_positions()[0]
    assert abs(p[2] - (0.1 + 0.005)) < 1e-3, p
