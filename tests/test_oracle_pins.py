"""Pins for the CPU restatement (oracle/liboracle.so): every stage checked
against something other than its own formula -- closed forms, textbook
routines, brute force, invariants, or the reference's own code compiled from
its sources (oracle/_ref).  The reference holds no golden vectors, KATs or
fixtures for this path, so these are the pins."""
import types

import numpy as np
import pytest

import oracle_lib as O
from vkxpbd import scenes

NUM_CELLS = 128 ** 3


def cfg(**kw):
    """A SolverConfig-shaped object with the reference defaults (scene.h:14-41)."""
    c = types.SimpleNamespace(particle_v_max=1e5, soft_contact_mu=0.5, shape_material_mu=1.0,
                              soft_contact_relaxation=0.9, particle_adhesion=0.0, particle_mu=0.5,
                              particle_cohesion=0.0, particle_max_radius=0.01, search_radius=0.02,
                              soft_contact_margin=0.01, soft_contact_restitution=0.0, ground_plane_altitude=0.0,
                              gravity=[0.0, 0.0, -9.81], enable_restitution=False, has_ground_plane=True)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def solver(q, r=0.01, qd=None, w=10.0, flags=None, fp64=False, meshes=(), **kw):
    q = np.asarray(q, np.float32).reshape(-1, 3)
    n = q.shape[0]
    o = O.OracleSolver(fp64=fp64)
    o.init(cfg(**kw), n)
    for v, t in meshes:
        o.add_mesh(v, t)
    o.finalize_meshes()
    wv = np.full(n, w, np.float32) if np.isscalar(w) else np.asarray(w, np.float32)
    o.set_particles(q, qd, wv, np.full(n, r, np.float32),
                    np.ones(n, np.int32) if flags is None else np.asarray(flags, np.int32))
    return o


def np_cells(q, inv):
    g = np.floor(q.astype(np.float32) * np.float32(inv)).astype(np.int64) + (1 << 20)
    g = np.maximum(g, 0) % 128
    return (g[:, 2] * 16384 + g[:, 1] * 128 + g[:, 0]).astype(np.uint32)


# ---------------------------------------------------------------- grid -----
def test_cell_index_closed_form():
    """grid_cell_indices.comp:17-25 on hand-computed points (cell 0.02 m)."""
    pts = [(0.0, 0.0, 0.0), (-1e-4, 0.0, 0.0), (0.025, 0.045, 0.065), (2.565, 0.0, 0.0), (-30000.0, 0.0, 0.0),
           (0.0, -0.0201, 0.0)]
    want = [0, 127, 3 * 16384 + 2 * 128 + 1, 0, 0, 126 * 128]
    o = solver(np.array(pts) + [0, 0, 5.0], gravity=[0, 0, 0], has_ground_plane=False)
    o.step(1.0 / 60.0, 1, 0)
    keys, ids = o.get_sorted()
    got = np.empty(len(pts), np.uint32)
    got[ids] = keys
    base = np.array(want, np.int64) + 250 * 16384   # z = 5.0 -> floor(250) + 2^20 -> 250 % 128 = 122?
    zc = (np.floor(np.float32(5.0) * np.float32(50.0)).astype(np.int64) + (1 << 20)) % 128
    np.testing.assert_array_equal(got, (np.array(want, np.int64) + zc * 16384).astype(np.uint32))
    assert zc == (250 + (1 << 20)) % 128
    del base


def test_sort_is_stable_sort_by_cell():
    """The radix sort's contract: a stable sort of (cell, id) by cell, i.e.
    numpy's stable argsort; plus permutation / order invariants."""
    rng = np.random.default_rng(1)
    n = 20000
    q = rng.uniform(-3.0, 3.0, size=(n, 3)).astype(np.float32)   # wide: wrap-around aliasing
    q[: n // 2] = rng.uniform(0.0, 0.05, size=(n // 2, 3))      # dense: many per cell
    o = solver(q, gravity=[0, 0, 0], has_ground_plane=False)
    o.step(1.0 / 60.0, 1, 0)
    keys, ids = o.get_sorted()
    cells = np_cells(q, 1.0 / np.float32(0.02))
    order = np.argsort(cells, kind="stable")
    np.testing.assert_array_equal(ids, order.astype(np.uint32))
    np.testing.assert_array_equal(keys, cells[order])
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)


def test_cell_tables_closed_form():
    rng = np.random.default_rng(2)
    q = rng.uniform(0.0, 0.2, size=(3000, 3)).astype(np.float32)
    o = solver(q, gravity=[0, 0, 0], has_ground_plane=False)
    o.step(1.0 / 60.0, 1, 0)
    keys, _ = o.get_sorted()
    st, en = o.get_cell_tables()
    u, first, cnt = np.unique(keys, return_index=True, return_counts=True)
    ws, we = np.zeros(NUM_CELLS, np.int32), np.zeros(NUM_CELLS, np.int32)
    ws[u], we[u] = first, first + cnt
    np.testing.assert_array_equal(st, ws)
    np.testing.assert_array_equal(en, we)
    assert int(np.sum(en - st)) == q.shape[0]


# ------------------------------------------------------- plane contacts ----
def test_plane_contact_set():
    """generate_plane_contacts.comp:27-54: {i active : z_i - alt < margin + r}."""
    rng = np.random.default_rng(3)
    n = 4000
    q = rng.uniform([-1, -1, -0.05], [1, 1, 0.08], size=(n, 3)).astype(np.float32)
    flags = (rng.uniform(size=n) > 0.1).astype(np.int32)
    o = solver(q, flags=flags, ground_plane_altitude=0.01)
    o.collide_only()
    c = o.get_contacts()
    want = np.nonzero((flags == 1) & ((q[:, 2] - np.float32(0.01)) < np.float32(0.01) + np.float32(0.01)))[0]
    np.testing.assert_array_equal(np.sort(c["particles"]), want)
    k = np.argsort(c["particles"])
    np.testing.assert_array_equal(c["shapes"], 0)
    np.testing.assert_array_equal(c["body_positions"][k][:, :2], q[want, :2])
    np.testing.assert_array_equal(c["body_positions"][:, 2], np.float32(0.01))
    np.testing.assert_array_equal(c["normals"], np.tile([0.0, 0.0, 1.0], (want.size, 1)))
    np.testing.assert_array_equal(c["body_velocities"], 0.0)


# ------------------------------------------------------------ integrate ----
@pytest.mark.parametrize("fp64,tol", [(True, 1e-12), (False, 2e-5)])
def test_free_fall_closed_form(fp64, tol):
    """No contacts: x_k = x_0 + v_0 k h + g h^2 k(k+1)/2, v_k = v_0 + k g h."""
    g = np.array([0.3, -1.0, -9.81])
    q = np.array([[0.0, 0.0, 10.0], [1.0, 2.0, 20.0], [-3.0, 0.5, 15.0]])
    v0 = np.array([[0.0, 0.0, 0.0], [0.5, -0.25, 1.0], [-2.0, 0.0, 3.0]], np.float32)
    o = solver(q, qd=v0, fp64=fp64, gravity=list(g), has_ground_plane=False)
    dt, S = 1.0 / 60.0, 8
    o.step(dt, S, 1)
    h = np.float64(np.float32(dt) / np.float32(S)) if not fp64 else np.float64(np.float32(np.float32(dt) / S))
    k = S
    x = q + v0 * k * h + g * h * h * k * (k + 1) / 2
    v = v0 + g * k * h
    np.testing.assert_allclose(o.get_positions(), x, atol=tol * 30)
    np.testing.assert_allclose(o.get_velocities(), v, atol=tol * 1e3)


def test_velocity_clamp():
    """|v| above particle_v_max is scaled back onto the sphere |v| = v_max."""
    o = solver([[0.0, 0.0, 10.0]], qd=[[300.0, 400.0, 0.0]], particle_v_max=10.0, gravity=[0, 0, 0],
               has_ground_plane=False)
    o.step(1.0 / 60.0, 1, 1)
    v = o.get_velocities()[0]
    np.testing.assert_allclose(np.linalg.norm(v), 10.0, rtol=1e-6)
    np.testing.assert_allclose(v / np.linalg.norm(v), [0.6, 0.8, 0.0], atol=1e-6)


# ----------------------------------------------------- particle-particle ----
@pytest.mark.parametrize("w1,w2", [(10.0, 10.0), (10.0, 2.5), (4.0, 0.0)])
def test_two_particle_head_on(w1, w2):
    """Overlap e along x, at rest: particle i moves by relax*e*w_i/(w1+w2)
    along its outward normal (solve_particle_particle_contacts.comp:150-177)."""
    r, e, relax = 0.01, 0.002, 0.9
    d = 2 * r - e
    q = [[-d / 2, 0.0, 1.0], [d / 2, 0.0, 1.0]]
    o = solver(q, r=r, w=[w1, w2], fp64=True, gravity=[0, 0, 0], has_ground_plane=False)
    h = np.float32(1.0 / 60.0)
    o.step(float(h), 1, 1)
    x = o.get_positions()
    dx0 = -(relax * e * w1 / (w1 + w2))
    dx1 = relax * e * w2 / (w1 + w2)
    np.testing.assert_allclose(x[0], [np.float32(-d / 2) + dx0, 0.0, 1.0], atol=1e-9)
    np.testing.assert_allclose(x[1], [np.float32(d / 2) + dx1, 0.0, 1.0], atol=1e-9)
    np.testing.assert_allclose(o.get_velocities()[:, 0], np.array([dx0, dx1]) / h, rtol=1e-6, atol=1e-9)


def test_pp_friction_cone():
    """Sliding pair: the tangential correction is capped at mu * |normal correction|
    and opposes the relative tangential velocity."""
    r, e, relax, mu = 0.01, 0.002, 0.9, 0.5
    d = 2 * r - e
    q = [[-d / 2, 0.0, 1.0], [d / 2, 0.0, 1.0]]
    qd = np.array([[0.0, 0.1, 0.0], [0.0, -0.1, 0.0]])
    o = solver(q, r=r, qd=qd, fp64=True, gravity=[0, 0, 0], has_ground_plane=False)
    h = float(np.float32(1.0 / 60.0))
    o.step(1.0 / 60.0, 1, 1)
    dl = o.get_deltas()
    xp = np.asarray(q, np.float32).astype(np.float64) + qd * h     # predicted positions
    n0 = (xp[0] - xp[1]) / np.linalg.norm(xp[0] - xp[1])
    dn = dl[0] @ n0
    dt_ = dl[0] - dn * n0
    assert dn > 0
    np.testing.assert_allclose(np.linalg.norm(dt_), mu * dn, rtol=1e-3)
    assert dt_[1] < 0   # particle 0 slides +y relative to particle 1: friction pushes -y


def test_pp_bruteforce_neighbour_set():
    """Grid query vs an O(N^2) neighbour search: the deltas equal the brute-force
    sum over all overlapping pairs (at rest: normal terms only), including
    particles aliased into the same cells 128 cells apart."""
    rng = np.random.default_rng(4)
    r = 0.01
    n = 400
    q = rng.uniform(0.0, 0.12, size=(n, 3))
    q[n // 2:] = q[: n // 2] + [128 * 0.02, 0.0, 0.0]   # same physical cells, far away
    q = q.astype(np.float32)
    w = rng.uniform(1.0, 10.0, size=n).astype(np.float32)
    o = solver(q, r=r, w=w, fp64=True, gravity=[0, 0, 0], has_ground_plane=False)
    o.step(1.0 / 60.0, 1, 1)
    got = o.get_deltas()
    X = q.astype(np.float64)
    diff = X[:, None, :] - X[None, :, :]
    dist = np.linalg.norm(diff, axis=2)
    rr = np.float64(np.float32(r))          # the solver's parameters are float32
    err = dist - 2 * rr
    mask = (err <= 0) & (dist > 0)
    nrm = np.where(dist[..., None] > 0, diff / np.maximum(dist, 1e-30)[..., None], 0.0)
    W = w.astype(np.float64)
    terms = -nrm * err[..., None] / (W[:, None] + W[None, :])[..., None]
    want = np.float64(np.float32(0.9)) * W[:, None] * np.sum(np.where(mask[..., None], terms, 0.0), axis=1)
    assert mask.sum() > 100
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-15)


# ------------------------------------------------------- shape contacts ----
def test_particle_pressed_into_plane():
    """c = z - r = -e: the correction lifts by relax*e (w2 = 0 for the static
    plane); with fast sliding the friction cap is mu_avg*e*relax backwards,
    mu_avg = (soft_contact_mu + shape_material_mu)/2."""
    r, e, relax = 0.01, 0.003, 0.9
    o = solver([[0.0, 0.0, r - e]], r=r, fp64=True, gravity=[0, 0, 0])
    o.step(1.0 / 60.0, 1, 1)
    np.testing.assert_allclose(o.get_deltas()[0], [0, 0, relax * e], atol=1e-9)
    o = solver([[0.0, 0.0, r - e]], r=r, qd=[[3.0, 0.0, 0.0]], fp64=True, gravity=[0, 0, 0])
    o.step(1.0 / 60.0, 1, 1)
    np.testing.assert_allclose(o.get_deltas()[0], [-0.75 * e * relax, 0, relax * e], atol=1e-9)


def test_resting_particle_has_no_drift():
    """A particle exactly touching the plane (c = 0), no gravity: nothing moves."""
    o = solver([[0.3, -0.2, 0.01]], r=0.01, gravity=[0, 0, 0])
    for _ in range(10):
        o.step(1.0 / 60.0, 8, 1)
    np.testing.assert_array_equal(o.get_positions()[0], np.float32([0.3, -0.2, 0.01]))


# ------------------------------------------------------- mesh narrowphase ---
def test_woop_ray_triangle():
    a, b, c = [0.0, -1.0, -1.0], [0.0, 1.0, -1.0], [0.0, 0.0, 1.0]
    assert O.ray_tri([-1.0, 0.0, 0.0], [1.0, 0.0, 0.0], a, b, c)
    assert not O.ray_tri([1.0, 0.0, 0.0], [1.0, 0.0, 0.0], a, b, c)          # behind
    assert not O.ray_tri([-1.0, 5.0, 0.0], [1.0, 0.0, 0.0], a, b, c)         # beside
    assert O.ray_tri([-1.0, 0.0, 0.0], [1.0, 0.01, -0.02], a, b, c)


def test_closest_point_triangle_bruteforce():
    """Closest point vs dense barycentric sampling (an upper bound it must
    meet) and barycentrics that reproduce it."""
    rng = np.random.default_rng(6)
    uv = np.array([(i / 200, j / 200) for i in range(201) for j in range(201 - i)])
    for _ in range(50):
        a, b, c = rng.normal(size=(3, 3)).astype(np.float32)
        p = rng.normal(size=3).astype(np.float32) * 1.5
        cp, bc = O.closest_point_triangle(p, a, b, c)
        s = a + uv[:, :1] * (b - a) + uv[:, 1:] * (c - a)
        best = np.min(np.linalg.norm(s - p, axis=1))
        d = np.linalg.norm(cp.astype(np.float64) - p)
        assert d <= best + 1e-5
        np.testing.assert_allclose(bc[0] * a + bc[1] * b + bc[2] * c, cp, atol=1e-5)
        np.testing.assert_allclose(bc.sum(), 1.0, atol=1e-6)


def test_mesh_contacts_sphere_and_reference_bruteforce():
    """Icosphere narrowphase: body_pos is the closest surface point (checked
    against the reference's own brute-force query, mesh_bvh.h:296-321), the
    normal points outward and the inside/outside sign matches the analytic
    sphere (inscribed / circumscribed radius)."""
    verts, tris = scenes.icosphere(2)
    rng = np.random.default_rng(8)
    dirs = rng.normal(size=(600, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    rad = 0.5 + rng.uniform(-0.03, 0.03, size=(600, 1))
    q = (dirs * rad + [0.0, 0.0, 0.5]).astype(np.float32)
    r = 0.005
    o = solver(q, r=r, meshes=[(verts, tris)], has_ground_plane=False, gravity=[0, 0, 0])
    o.collide_only()
    c = o.get_contacts()
    assert c["particles"].size > 100
    cen = np.array([0.0, 0.0, 0.5])
    assert np.all(np.sum((c["body_positions"] - cen) * c["normals"], axis=1) > 0)
    # inscribed radius of the level-2 icosphere: minimum face-plane distance
    a, b, cc = (verts[tris[:, k]].astype(np.float64) - cen for k in range(3))
    nf = np.cross(b - a, cc - a)
    r_in = np.min(np.abs(np.sum(nf * a, 1)) / np.linalg.norm(nf, axis=1))
    for pi, bp, nn in zip(c["particles"], c["body_positions"], c["normals"]):
        p = q[pi].astype(np.float64)
        dc = np.linalg.norm(p - cen)
        signed = np.dot(p - bp, nn)
        if dc < r_in - 1e-6:
            assert signed < 0          # inside: penetration
        if dc > 0.5 + 1e-6:
            assert signed > 0          # outside the circumscribed sphere
        if O.ref_lib() is not None:
            ok, face, pt = O.ref_query_closest(verts, tris, q[pi], 0.5)
            assert ok
            # same distance; the point itself may differ where two faces tie
            np.testing.assert_allclose(np.linalg.norm(p - bp), np.linalg.norm(p - pt), atol=1e-7)
            np.testing.assert_allclose(bp, pt, atol=1e-4)


def test_fp32_restatement_tracks_fp64():
    """The fp32 step against the FP64 host step on a small pile: after one
    frame the absolute error is far below the 1e-5 m bar.  A granular pile is
    chaotic, so over many frames fp32 and fp64 trajectories separate; that
    drift is reported, not bounded."""
    sc = scenes.pile_4096()
    p = sc["particles"]
    sub = slice(0, 1024)
    outs = {}
    for fp64 in (False, True):
        o = solver(p["q"][sub], r=sc["r"], fp64=fp64)
        o.step(1.0 / 60.0, 8, 1)
        outs[(fp64, 1)] = o.get_positions()
        for _ in range(9):
            o.step(1.0 / 60.0, 8, 1)
        outs[(fp64, 10)] = o.get_positions()
    err1 = float(np.max(np.abs(outs[(False, 1)] - outs[(True, 1)])))
    err10 = float(np.max(np.abs(outs[(False, 10)] - outs[(True, 10)])))
    print(f"fp32 vs fp64: 1 frame {err1:.3e} m, 10 frames {err10:.3e} m (drift)")
    assert err1 < 1e-6, err1


# ------------------------------------------------------------ restitution ---
def _one_particle(z, vz, vx=0.0, e=0.5, restitution=True, **kw):
    o = solver([[0.1, -0.2, z]], r=0.01, qd=[[vx, 0.0, vz]], fp64=True, gravity=[0, 0, 0],
               enable_restitution=restitution, soft_contact_restitution=e, **kw)
    o.step(1.0 / 60.0, 1, 1)
    return o.get_positions()[0], o.get_velocities()[0]


@pytest.mark.parametrize("e", [0.0, 0.5, 1.0])
def test_restitution_approaching_particle(e):
    """apply_particle_shape_restitution.comp:48-116: a particle in contact at
    the substep start (c = z - r <= adhesion) that approaches the plane with
    normal speed v_n leaves with exactly -e*v_n along the normal, whatever the
    position solve left; tangential velocity and position are the
    restitution-free step's."""
    vz, vx = -1.0, 0.3
    x, v = _one_particle(0.01 - 0.001, vz, vx, e)
    x0, v0 = _one_particle(0.01 - 0.001, vz, vx, e, restitution=False)
    np.testing.assert_allclose(v[2], -e * vz, atol=1e-12)
    assert v0[2] != pytest.approx(-e * vz, abs=1e-3)          # the solve alone gives something else
    np.testing.assert_array_equal(x, x0)
    np.testing.assert_array_equal(v[:2], v0[:2])


def test_restitution_uses_substep_start_velocity():
    """The approach speed is the velocity before integration (v_init,
    integrate_particles.comp:34-35), not the predicted one: with gravity g the
    result is -e*v0_z, not -e*(v0_z + g h)."""
    e, vz = 0.5, -1.0
    o = solver([[0.0, 0.0, 0.009]], r=0.01, qd=[[0.0, 0.0, vz]], fp64=True, gravity=[0, 0, -9.81],
               enable_restitution=True, soft_contact_restitution=e)
    o.step(1.0 / 60.0, 1, 1)
    np.testing.assert_allclose(o.get_velocities()[0, 2], -e * vz, atol=1e-12)


@pytest.mark.parametrize("z,vz", [(0.009, 0.2), (0.012, -1.0)])
def test_restitution_leaves_receding_and_separated_particles(z, vz):
    """A receding particle (v_n > 0) and one whose substep-start gap
    c = z - r is positive (above the adhesion distance 0) are untouched: the
    state equals the restitution-free step bit for bit."""
    x, v = _one_particle(z, vz)
    x0, v0 = _one_particle(z, vz, restitution=False)
    np.testing.assert_array_equal(x, x0)
    np.testing.assert_array_equal(v, v0)


def test_restitution_relative_to_moving_mesh():
    """The contact's body velocity enters the relative normal velocity: a flat
    mesh quad moving up at u leaves the particle with u + e*(u - v0) along
    the normal (the mesh vertex velocities, interpolated at the contact)."""
    e, vz, u = 0.5, -1.0, 0.4
    verts = np.float32([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]])
    tris = np.uint32([[0, 1, 2], [0, 2, 3]])
    o = solver([[0.1, 0.2, 0.009]], r=0.01, qd=[[0.0, 0.0, vz]], fp64=True, gravity=[0, 0, 0],
               meshes=[(verts, tris)], has_ground_plane=False, enable_restitution=True,
               soft_contact_restitution=e)
    o.update_mesh(0, verts, np.tile(np.float32([0.0, 0.0, u]), (4, 1)))
    o.step(1.0 / 60.0, 1, 1)
    assert o.last_contact_count == 1
    np.testing.assert_allclose(o.get_velocities()[0, 2], u + e * (u - vz), atol=1e-9)


# ---------------------------------------------------- cohesion / adhesion ---
@pytest.mark.parametrize("w1,w2", [(10.0, 10.0), (10.0, 2.5)])
def test_cohesion_attracts_separated_pair(w1, w2):
    """solve_particle_particle_contacts.comp:150-177 with cohesion: a pair at
    rest with gap 0 < d - 2r <= cohesion is a contact; each particle moves
    toward the other by relax*(d - 2r)*w_i/(w1 + w2).  A pair whose gap
    exceeds the cohesion distance does not move."""
    r, coh, relax = 0.01, 0.002, 0.9
    for gap, touch in ((0.5 * coh, True), (1.5 * coh, False)):
        d = 2 * r + gap
        q = np.float32([[-d / 2, 0.0, 1.0], [d / 2, 0.0, 1.0]])
        o = solver(q, r=r, w=[w1, w2], fp64=True, gravity=[0, 0, 0], has_ground_plane=False,
                   particle_cohesion=coh)
        o.step(1.0 / 60.0, 1, 1)
        x = o.get_positions()
        err = float(q[1, 0]) - float(q[0, 0]) - 2 * float(np.float32(r))
        if touch:
            np.testing.assert_allclose(x[0, 0] - q[0, 0], relax * err * w1 / (w1 + w2), rtol=1e-6)
            np.testing.assert_allclose(x[1, 0] - q[1, 0], -relax * err * w2 / (w1 + w2), rtol=1e-6)
        else:
            np.testing.assert_array_equal(x, q)
        np.testing.assert_array_equal(x[:, 1:], q[:, 1:])


def test_adhesion_pulls_particle_to_plane():
    """solve_particle_shape_contacts.comp with particle_adhesion as the contact
    distance ka: a particle hovering c = z - r above the plane with c <= ka
    is pulled down by relax*c; with c > ka it stays."""
    r, c, relax = 0.01, 0.001, 0.9
    for ka, moves in ((0.002, True), (0.0005, False)):
        o = solver([[0.0, 0.0, r + c]], r=r, fp64=True, gravity=[0, 0, 0], particle_adhesion=ka)
        o.step(1.0 / 60.0, 1, 1)
        dz = o.get_deltas()[0]
        want = -relax * (float(np.float32(r + c)) - float(np.float32(r))) if moves else 0.0
        np.testing.assert_allclose(dz, [0.0, 0.0, want], atol=1e-12)


# ------------------------------------------- non-watertight mesh normals ---
def _open_fan(ring_angles, h, rad=0.1, apex=(0.0, 0.0, 0.5)):
    """An open cone: apex vertex 0 and a ring of vertices h below it, one
    triangle per ring edge wound so the normals point up and out; the
    bottom is left open (not watertight)."""
    ring = np.stack([rad * np.cos(ring_angles), rad * np.sin(ring_angles), np.full_like(ring_angles, -h)], 1)
    verts = (np.concatenate([np.zeros((1, 3)), ring]) + np.asarray(apex)).astype(np.float32)
    k = len(ring_angles)
    tris = np.array([[0, 1 + j, 1 + (j + 1) % k] for j in range(k)], np.uint32)
    return verts, tris


def _fan_normals(verts, tris):
    """Angle-weighted (apex angle, in float64 from arccos) and uniform sums of
    the unit face normals around the apex."""
    v = verts.astype(np.float64)
    nt, n1 = np.zeros(3), np.zeros(3)
    for t in tris:
        a, b, c = v[t]
        n = np.cross(b - a, c - a)
        n /= np.linalg.norm(n)
        e0, e1 = (b - a) / np.linalg.norm(b - a), (c - a) / np.linalg.norm(c - a)
        nt += np.arccos(np.clip(e0 @ e1, -1.0, 1.0)) * n
        n1 += n
    return nt, n1


def test_open_mesh_face_interior_normal():
    """generate_mesh_contacts.comp:266-313,390 on an open mesh: a point over a
    face interior gets that face's normal, on both sides of the face; the
    signed distance is negative behind it and body_pos is the foot point."""
    verts, tris = _open_fan(np.array([0.0, 2.1, 4.2]), 0.02)
    v = verts.astype(np.float64)
    a, b, c = v[tris[0]]
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n)
    foot = (a + b + c) / 3.0
    r, t = 0.005, 0.008
    q = np.float32([foot + t * n, foot - t * n])
    o = solver(q, r=r, meshes=[(verts, tris)], fp64=True, has_ground_plane=False, gravity=[0, 0, 0])
    o.collide_only()
    cts = o.get_contacts()
    assert sorted(cts["particles"].tolist()) == [0, 1]
    for k, pi in enumerate(cts["particles"]):
        np.testing.assert_allclose(cts["normals"][k], n, atol=1e-6)
        np.testing.assert_allclose(cts["body_positions"][k], foot, atol=1e-6)
        sd = float(np.dot(q[pi].astype(np.float64) - cts["body_positions"][k], cts["normals"][k]))
        assert (sd > 0) == (pi == 0)


@pytest.mark.parametrize("side,h", [(1.0, 0.3), (-1.0, -0.3)])
def test_open_mesh_vertex_region_uses_angle_weights(side, h):
    """A point in the apex's region of an open, uneven cone: every face's
    closest point is the apex, and the inside/outside sign is the sign of the
    angle-weighted pseudo-normal (weights = apex angles, :276-281) dotted
    with p - apex.  The direction is chosen where the uniform sum of face
    normals gives the opposite sign, so a wrong weight fails the test.  A
    spike (h > 0) has only outside points in its apex region, a pit (ring
    above the apex, h < 0) only inside points."""
    verts, tris = _open_fan(np.array([0.0, 0.25, 0.5, 2.6, 4.4]), h)
    nt, n1 = _fan_normals(verts, tris)
    apex = verts[0].astype(np.float64)
    rel = verts[1:].astype(np.float64) - apex
    rng = np.random.default_rng(12)
    u = None
    for _ in range(100000):
        w = rng.normal(size=3)
        w /= np.linalg.norm(w)
        if (np.all(rel @ w < -1e-2) and np.sign(w @ nt) == side and np.sign(w @ n1) == -side
                and min(abs(w @ nt), abs(w @ n1)) > 0.02):
            u = w
            break
    assert u is not None
    r, t = 0.005, 0.008
    q = np.float32([apex + t * u])
    o = solver(q, r=r, meshes=[(verts, tris)], fp64=True, has_ground_plane=False, gravity=[0, 0, 0])
    o.collide_only()
    cts = o.get_contacts()
    assert cts["particles"].tolist() == [0]
    np.testing.assert_allclose(cts["body_positions"][0], apex, atol=1e-9)
    dq = q[0].astype(np.float64) - apex   # the stored (float32) point's direction
    np.testing.assert_allclose(cts["normals"][0], side * dq / np.linalg.norm(dq), atol=1e-9)
