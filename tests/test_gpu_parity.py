"""HIP solver vs the oracle on identical inputs (through the vkxpbd module,
which calls the C-ABI of libvkxpbd.so).

Bar: integer stages (cell keys, sorted ids, cell tables, contact sets) bit
exact; floating-point state bit exact against the fp32 restatement wherever
the reference's float-atomic order cannot matter (at most two shape contacts
per particle), and always within 1e-5 m of it."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import assert_contacts_equal, gpu_contacts, log, make_pair, scene_from
from vkxpbd import scenes

pytestmark = pytest.mark.gpu

TOL_M = 1e-5


def compare_state(g, o, exact=True, what=""):
    gp, op = g.get_positions().astype(np.float64), o.get_positions()
    gv, ov = g.get_velocities().astype(np.float64), o.get_velocities()
    dx = float(np.max(np.abs(gp - op))) if gp.size else 0.0
    dv = float(np.max(np.abs(gv - ov))) if gv.size else 0.0
    nbad = int(np.count_nonzero(gp != op))
    log(f"  {what}: max|dx|={dx:.3e} max|dv|={dv:.3e} differing coords={nbad}")
    assert dx <= TOL_M
    if exact:
        np.testing.assert_array_equal(gp, op)
        np.testing.assert_array_equal(gv, ov)
    return dx


def compare_grid(g, o):
    gk, gi = g.get_sorted_cells(), g.get_sorted_point_ids()
    ok, oi = o.get_sorted()
    np.testing.assert_array_equal(gk, ok)
    np.testing.assert_array_equal(gi, oi)
    os_, oe = o.get_cell_tables()
    np.testing.assert_array_equal(g.get_cell_starts(), os_)
    np.testing.assert_array_equal(g.get_cell_ends(), oe)


def test_one_substep_bitexact_pile_4096(vk):
    sc = scenes.pile_4096()
    g, o = make_pair(vk, sc)
    log("pile_4096: step(1/60, 1, 1)")
    g.step(1.0 / 60.0, 1, 1)
    o.step(1.0 / 60.0, 1, 1)
    compare_grid(g, o)
    assert g.last_contact_count == o.last_contact_count
    assert_contacts_equal(gpu_contacts(g), o.get_contacts())
    np.testing.assert_array_equal(g.get_deltas().astype(np.float64), o.get_deltas())
    compare_state(g, o, exact=True, what="1 substep")


def test_frames_pile_4096(vk):
    """The 4,096-particle pile for 30 frames x 8 substeps, compared every frame."""
    sc = scenes.pile_4096()
    g, o = make_pair(vk, sc)
    for f in range(30):
        g.step(1.0 / 60.0, 8, 1)
        o.step(1.0 / 60.0, 8, 1)
        assert g.last_contact_count == o.last_contact_count
        if f % 10 == 9:
            compare_state(g, o, exact=True, what=f"frame {f + 1}")
            compare_grid(g, o)
    assert_contacts_equal(gpu_contacts(g), o.get_contacts())


@pytest.mark.parametrize("n", [1, 2, 63, 1023, 1025, 5000])
def test_ragged_sizes(vk, n):
    rng = np.random.default_rng(n)
    r = 0.01
    q = rng.uniform([-0.1, -0.1, 0.0], [0.1, 0.1, 0.15], size=(n, 3)).astype(np.float32)
    qd = rng.normal(0, 0.3, size=(n, 3)).astype(np.float32)
    sc = scene_from(q, r, qd)
    g, o = make_pair(vk, sc)
    log(f"ragged n={n}")
    for _ in range(3):
        g.step(1.0 / 60.0, 4, 1)
        o.step(1.0 / 60.0, 4, 1)
    compare_grid(g, o)
    compare_state(g, o, exact=True, what=f"n={n}")


@pytest.mark.parametrize("n", [1, 63, 1025])
def test_ragged_sizes_with_mesh(vk, n):
    """Particle counts that leave partial workgroups in every mesh pass (the
    four-particle candidate threads, the walk's per-XCD queue ranges) over a
    watertight icosphere, against the oracle bit for bit."""
    rng = np.random.default_rng(100 + n)
    r = 0.01
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    q = (np.array([0.0, 0.0, 0.5]) + u * rng.uniform(0.49, 0.53, size=(n, 1))).astype(np.float32)
    verts, tris = scenes.icosphere(3)
    sc = scene_from(q, r, meshes=[(verts, tris)])
    g, o = make_pair(vk, sc)
    log(f"ragged n={n} over an icosphere")
    for f in range(3):
        g.step(1.0 / 60.0, 4, 1)
        o.step(1.0 / 60.0, 4, 1)
        assert g.last_contact_count == o.last_contact_count
    assert g.last_contact_count > 0
    assert_contacts_equal(gpu_contacts(g), o.get_contacts())
    compare_state(g, o, exact=True, what=f"n={n} with mesh")


def test_two_iterations_cohesion_restitution(vk):
    rng = np.random.default_rng(11)
    r = 0.01
    q = scenes.lattice(12, 12, 6, 2 * r * 0.95, (-0.11, -0.11, 0.004), 0.1 * r, rng)
    qd = rng.normal(0, 0.2, size=q.shape).astype(np.float32)
    qd[:, 2] -= 1.0
    sc = scene_from(q, r, qd)
    g, o = make_pair(vk, sc, particle_cohesion=0.002, particle_adhesion=0.001, enable_restitution=True,
                     soft_contact_restitution=0.5)
    for k in range(5):
        g.step(1.0 / 60.0, 4, 2)
        o.step(1.0 / 60.0, 4, 2)
    # restitution updates velocities with float atomics on the velocity buffer
    # it also reads (apply_particle_shape_restitution.comp:112-114): with one
    # shape contact per particle the result is order independent
    compare_state(g, o, exact=True, what="K=2 + cohesion + restitution")


def test_inactive_and_proxy_particles(vk):
    rng = np.random.default_rng(5)
    r = 0.01
    q = scenes.lattice(10, 10, 4, 2 * r * 0.9, (-0.09, -0.09, 0.005), 0.1 * r, rng)
    flags = np.ones(q.shape[0], np.int32)
    flags[::7] = 0          # inactive
    flags[3::11] = 1 | 2    # active proxy
    sc = scene_from(q, r, flags=flags)
    g, o = make_pair(vk, sc)
    for _ in range(4):
        g.step(1.0 / 60.0, 8, 1)
        o.step(1.0 / 60.0, 8, 1)
    compare_state(g, o, exact=True, what="inactive/proxy")
    # inactive particles never move (the build's documented passthrough)
    off = flags == 0
    np.testing.assert_array_equal(g.get_positions()[off], q[off])


def test_collide_only_and_injected_contacts(vk):
    sc = scenes.pile_4096()
    g, o = make_pair(vk, sc)
    for _ in range(20):
        g.step(1.0 / 60.0, 8, 1)
        o.step(1.0 / 60.0, 8, 1)
    g.collide_only()
    o.collide_only()
    assert g.last_contact_count == o.last_contact_count > 0
    c = o.get_contacts()
    assert_contacts_equal(gpu_contacts(g), c)
    # inject a permuted half of the contact set and solve without collision
    rng = np.random.default_rng(0)
    sel = rng.permutation(c["particles"].size)[: c["particles"].size // 2]
    args = (c["particles"][sel].astype(np.int32), c["shapes"][sel].astype(np.int32),
            c["body_positions"][sel].astype(np.float32), c["normals"][sel].astype(np.float32), int(sel.size))
    g.set_contact_data(*args)
    o.set_contact_data(*args)
    assert g.last_contact_count == sel.size
    g.step_solve_only(1.0 / 60.0, 4, 1)
    o.step_solve_only(1.0 / 60.0, 4, 1)
    compare_state(g, o, exact=True, what="step_solve_only")


def ico_shell(level, n, seed, band=0.02, r=0.005):
    """Particles in a band around the icosphere surface, inside and outside."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rad = 0.5 + rng.uniform(-band, band, size=(n, 1))
    return (d * rad + [0.0, 0.0, 0.5]).astype(np.float32), r


@pytest.mark.parametrize("watertight", [True, False])
def test_mesh_contacts_collide_only(vk, watertight):
    verts, tris = scenes.icosphere(3)
    q, r = ico_shell(3, 4000, seed=1)
    sc = scene_from(q, r, meshes=[(verts, tris)])
    cfg = scenes.solver_config(vk, r)
    g = vk.XpbdSolver()
    g.init(cfg, q.shape[0])
    g.add_mesh(verts.reshape(-1), tris.reshape(-1), is_watertight=watertight)
    g.finalize_meshes()
    p = sc["particles"]
    g.set_particles(p["q"], p["qd"], p["inv_mass"], p["radii"], p["flags"])
    o = O.OracleSolver()
    o.init(cfg, q.shape[0])
    o.add_mesh(verts, tris, is_watertight=watertight)
    o.finalize_meshes()
    o.set_particles(p["q"], p["qd"], p["inv_mass"], p["radii"], p["flags"])
    g.collide_only()
    o.collide_only()
    log(f"mesh collide_only watertight={watertight}: contacts gpu={g.last_contact_count} oracle={o.last_contact_count}")
    assert g.last_contact_count == o.last_contact_count > 0
    gc, oc = gpu_contacts(g), o.get_contacts()
    if watertight:
        assert_contacts_equal(gc, oc)
    else:
        # the pseudo-normal weights use acos (device vs host libm): sets and
        # positions must agree, normals to a tolerance
        a, b = (lambda c: {k: np.asarray(v)[np.lexsort((c["shapes"], c["particles"]))] for k, v in c.items()})(gc), \
               (lambda c: {k: np.asarray(v)[np.lexsort((c["shapes"], c["particles"]))] for k, v in c.items()})(oc)
        np.testing.assert_array_equal(a["particles"], b["particles"])
        np.testing.assert_allclose(a["normals"], b["normals"], atol=1e-6)
    # outward normals: the mesh contact's normal points away from the centre
    m = gc["shapes"] == 1
    rel = gc["body_positions"][m] - [0.0, 0.0, 0.5]
    assert np.all(np.sum(rel * gc["normals"][m], axis=1) > 0.0)


def test_mesh_contacts_several_meshes_and_large_particles(vk):
    """Two watertight spheres (one small and offset, with a shape margin) and an
    empty mesh; a tenth of the particles are 6x larger than
    particle_max_radius, which takes their search beyond the distance-bound
    grid's cap (the search then starts at the root).  Contacts equal the
    oracle's brute-force narrowphase bit for bit, except for the empty mesh:
    the reference points an empty mesh's root at the next mesh's node 0
    (scene.cpp:291, root = node_base + 0 with no nodes of its own), so it and
    the oracle report phantom contacts with that leaf's triangle under the
    empty mesh's shape; this build reports none (a deliberate deviation)."""
    v1, t1 = scenes.icosphere(3)
    v2, t2 = scenes.icosphere(2, 0.2, (0.45, 0.1, 0.95))
    q, r = ico_shell(3, 3000, seed=5, band=0.04)
    rng = np.random.default_rng(6)
    d = rng.normal(size=(1500, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    q2 = (d * (0.2 + rng.uniform(-0.03, 0.03, (1500, 1))) + [0.45, 0.1, 0.95]).astype(np.float32)
    q = np.concatenate([q, q2])
    radii = np.full(q.shape[0], r, np.float32)
    radii[::10] = 6 * r
    empty_v, empty_t = np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32)
    meshes = [(v1, t1, 0.0), (empty_v, empty_t, 0.0), (v2, t2, 0.004)]
    cfg = scenes.solver_config(vk, r)
    g = vk.XpbdSolver()
    g.init(cfg, q.shape[0])
    o = O.OracleSolver()
    o.init(cfg, q.shape[0])
    for v, t, m in meshes:
        g.add_mesh(v.reshape(-1), t.reshape(-1), shape_margin=m)
        o.add_mesh(v, t, shape_margin=m)
    g.finalize_meshes()
    o.finalize_meshes()
    n = q.shape[0]
    args = (q, np.zeros_like(q), np.full(n, scenes.PARTICLE_INV_MASS, np.float32), radii, np.ones(n, np.int32))
    g.set_particles(*args)
    o.set_particles(*args)
    g.collide_only()
    o.collide_only()
    gc, oc = gpu_contacts(g), o.get_contacts()
    log(f"two meshes + empty mesh, large particles: contacts gpu={g.last_contact_count} oracle={o.last_contact_count}")
    assert not np.any(gc["shapes"] == 2)
    phantom = np.asarray(oc["shapes"]) == 2
    log(f"  oracle phantom contacts under the empty mesh: {int(np.count_nonzero(phantom))}")
    oc = {k: np.asarray(v)[~phantom] for k, v in oc.items()}
    assert g.last_contact_count == len(oc["shapes"]) > 0
    assert set(np.unique(gc["shapes"])) >= {1, 3}
    big = np.isin(gc["particles"], np.arange(0, n, 10))
    assert np.count_nonzero(big) > 0
    assert_contacts_equal(gc, oc)


def test_two_meshes_stepped(vk):
    """A pile dropped onto two intersecting spheres, away from the ground
    plane: particles near the crease get one contact from each mesh, solved
    by each mesh's walk (two float atomics into a zeroed sum, the same in
    either order), so the state equals the oracle's bit for bit."""
    rng = np.random.default_rng(51)
    r = 0.01
    v1, t1 = scenes.icosphere(3)
    v2, t2 = scenes.icosphere(2, 0.3, (0.35, 0.0, 0.8))
    q = scenes.lattice(16, 12, 4, 2 * r, (0.05, -0.11, 1.12), 0.1 * r, rng)
    sc = scene_from(q, r, meshes=[(v1, t1), (v2, t2)])
    g, o = make_pair(vk, sc)
    both = 0
    for f in range(12):
        g.step(1.0 / 60.0, 8, 1)
        o.step(1.0 / 60.0, 8, 1)
        assert g.last_contact_count == o.last_contact_count
        c = gpu_contacts(g)
        ps, sh = np.asarray(c["particles"]), np.asarray(c["shapes"])
        both = max(both, len(np.intersect1d(ps[sh == 1], ps[sh == 2])))
        if f % 4 == 3:
            compare_state(g, o, exact=True, what=f"two meshes, frame {f + 1}")
    log(f"  particles touching both meshes in one substep: up to {both}")
    assert both > 0
    assert_contacts_equal(gpu_contacts(g), o.get_contacts())


def test_profiled_frames_equal_graph_frames(vk):
    """Stage timing runs every kernel on one stream (mesh stage included, the
    walk on its full grid) with events between stages: the same bits as the
    two-stream graph frames, frame after frame."""
    rng = np.random.default_rng(3)
    r = 0.01
    q = scenes.lattice(24, 24, 8, 2 * r, (-0.23, -0.23, 1.03), 0.1 * r, rng)
    verts, tris = scenes.icosphere(3)
    sc = scene_from(q, r, meshes=[(verts, tris)])
    a = scenes.build_solver(vk, sc)
    b = scenes.build_solver(vk, sc, enable_profile=True)
    for f in range(16):
        a.step(1.0 / 60.0, 8, 1)
        b.step(1.0 / 60.0, 8, 1)
        np.testing.assert_array_equal(a.get_positions(), b.get_positions(), err_msg=f"frame {f}")
        np.testing.assert_array_equal(a.get_velocities(), b.get_velocities(), err_msg=f"frame {f}")
        assert a.last_contact_count == b.last_contact_count
    assert sum(n for _, n in b.profile_stages().values()) > 0
    log(f"profiled frames == graph frames over 16 frames, contacts {a.last_contact_count}")


def test_diagnostic_pair_records(vk):
    """enable_diagnostics: one record per contributing pair (scene.h:43-64).
    A floating, overlapping lattice (no shape contacts) after one substep:
    the records' (i, j) set is the brute-force set of pairs closer than
    r_i + r_j at the predicted positions, their distances agree with those
    positions, and each particle's records sum to its applied delta (the
    oracle's, bit for bit, through get_deltas)."""
    rng = np.random.default_rng(8)
    r = 0.01
    q = scenes.lattice(6, 6, 6, 1.9 * r, (-0.05, -0.05, 1.0), 0.05 * r, rng)
    n = q.shape[0]
    sc = scene_from(q, r)
    g, o = make_pair(vk, sc, enable_diagnostics=True, max_diag_pairs=100000)
    dt = 1.0 / 60.0
    g.step(dt, 1, 1)
    o.step(dt, 1, 1)
    compare_state(g, o, exact=True, what="diagnostics substep")
    np.testing.assert_array_equal(g.get_deltas().astype(np.float64), o.get_deltas())
    recs = g.get_diag_pairs()
    assert g.get_diag_overflow_count() == 0
    assert g.get_diag_pair_count() == len(recs) > 0
    # predicted positions (zero initial velocity, gravity only), in float64
    x1 = q.astype(np.float64) + np.array([0.0, 0.0, -9.81 * dt * dt])
    d = np.linalg.norm(x1[:, None, :] - x1[None, :, :], axis=2)
    err = d - 2 * r
    want = {(i, j) for i in range(n) for j in range(n) if i != j and err[i, j] < -1e-7}
    unsure = {(i, j) for i in range(n) for j in range(n) if i != j and abs(err[i, j]) <= 1e-7}
    got = {(rc["i"], rc["j"]) for rc in recs}
    assert len(got) == len(recs)                      # one record per ordered pair
    assert want <= got and got <= want | unsure
    acc = np.zeros((n, 3))
    for rc in recs:
        i, j = rc["i"], rc["j"]
        assert rc["iter_index"] == 0
        assert abs(rc["dist"] - d[i, j]) <= 1e-6 and abs(rc["err"] - err[i, j]) <= 1e-6
        acc[i] += rc["delta_i"]
    np.testing.assert_allclose(acc, g.get_deltas().astype(np.float64), rtol=1e-5, atol=1e-9)
    log(f"diagnostics: {len(recs)} pair records over {n} particles")
    g.reset_diag_buffers()
    assert g.get_diag_pair_count() == 0


@pytest.mark.parametrize("max_contacts", [0, 5999], ids=["plane_in_pair_solve", "plane_in_shape_solve"])
def test_particles_touching_plane_and_mesh(vk, max_contacts):
    """Particles in the wedge where a sphere meets the ground carry a plane
    and a mesh contact at once: their two shape deltas are summed and applied
    after the particle-pair solve by the owner mark (the larger contact
    index).  States, deltas and contacts equal the oracle's bit for bit over
    several substeps, and particles with both contacts are present.  With
    the default contact capacity (2N) the plane contacts are solved inside
    the particle-pair solve; with room for fewer than every possible contact
    (2N - 1) they go through the shape solve.  1,000 particles high above
    the scene keep the contacts emitted (at most two per wedge particle)
    below either capacity, so both runs must give the oracle's bits."""
    rng = np.random.default_rng(31)
    r = 0.005
    n = 2000
    ang = rng.uniform(0.0, 2.0 * np.pi, n)
    z = rng.uniform(0.004, 0.012, n)
    rho = np.sqrt(0.5 ** 2 - (0.5 - z) ** 2) + rng.uniform(0.0, 0.012, n)   # just outside the sphere
    q = np.stack([rho * np.cos(ang), rho * np.sin(ang), z], 1).astype(np.float32)
    far = scenes.lattice(10, 10, 10, 3 * r, (-0.05, -0.05, 3.0), 0.0, rng)   # nothing within reach
    q = np.concatenate([q, far]).astype(np.float32)
    verts, tris = scenes.icosphere(4)
    sc = scene_from(q, r, meshes=[(verts, tris)])
    g, o = make_pair(vk, sc, max_contacts=max_contacts)
    for f in range(3):
        g.step(1.0 / 60.0, 4, 1)
        o.step(1.0 / 60.0, 4, 1)
        assert g.last_contact_count == o.last_contact_count
        assert g.last_contact_count <= 2 * n   # every emitted contact was kept
        compare_state(g, o, exact=True, what=f"wedge frame {f + 1}")
    np.testing.assert_array_equal(g.get_deltas().astype(np.float64), o.get_deltas())
    c = gpu_contacts(g)
    assert_contacts_equal(c, o.get_contacts())
    plane = set(c["particles"][c["shapes"] == 0].tolist())
    mesh = set(c["particles"][c["shapes"] == 1].tolist())
    both = plane & mesh
    log(f"wedge: {len(plane)} plane, {len(mesh)} mesh contacts, {len(both)} particles with both")
    assert len(both) > 50


def test_mesh_pile_frames(vk):
    """A pile falling onto a watertight icosphere: the config-1 path at small size."""
    rng = np.random.default_rng(2)
    r = 0.01
    q = scenes.lattice(24, 24, 8, 2 * r, (-0.23, -0.23, 1.03), 0.1 * r, rng)
    verts, tris = scenes.icosphere(3)
    sc = scene_from(q, r, meshes=[(verts, tris)])
    g, o = make_pair(vk, sc)
    for f in range(24):
        g.step(1.0 / 60.0, 8, 1)
        o.step(1.0 / 60.0, 8, 1)
        assert g.last_contact_count == o.last_contact_count
        if f % 6 == 5:
            log(f"  frame {f + 1}: contacts={g.last_contact_count}")
            assert_contacts_equal(gpu_contacts(g), o.get_contacts())
            compare_state(g, o, exact=True, what=f"mesh pile frame {f + 1}")


@pytest.mark.parametrize("env", [{"VKXPBD_SORT_CHUNKED": "1"},
                                 {"VKXPBD_SORT_CHUNKED": "1", "VKXPBD_WRAP_PLANES": "1"},
                                 {"VKXPBD_FUSE_KEYS": "0"},
                                 {"VKXPBD_SOLVE_IN_WALK": "0"}],
                         ids=["chunked_radix", "chunked_radix_wrap_planes", "integrate_every_substep",
                              "shape_solve_launched"])
def test_sort_and_keying_variants(vk, env, monkeypatch):
    """The radix passes without a bin-scan launch (per-chunk totals, 65,536
    keys: two chunks of 32 tiles), with plain and wrap-plane keys, and
    substeps all keyed by k_integrate give the oracle's bits, as the default
    path (later substeps keyed by the previous substep's apply) does: a pile
    landing on an icosphere, compared every frame."""
    rng = np.random.default_rng(21)
    r = 0.005
    q = scenes.lattice(64, 64, 16, 2 * r, (-0.315, -0.315, 0.505), 0.1 * r, rng)
    verts, tris = scenes.icosphere(3)
    sc = scene_from(q, r, meshes=[(verts, tris)])
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    g, o = make_pair(vk, sc)
    for k in env:
        monkeypatch.delenv(k)
    d = scenes.build_solver(vk, sc)   # the defaults
    log(f"variant {env}: wrap planes {g.wrap_planes}")
    for f in range(4):
        for s in (g, o, d):
            s.step(1.0 / 60.0, 8, 1)
        assert g.last_contact_count == o.last_contact_count == d.last_contact_count
        compare_state(g, o, exact=True, what=f"variant frame {f + 1}")
        np.testing.assert_array_equal(g.get_positions(), d.get_positions())
        np.testing.assert_array_equal(g.get_velocities(), d.get_velocities())
    compare_grid(g, o)
    assert_contacts_equal(gpu_contacts(g), o.get_contacts())


@pytest.mark.parametrize("iterations", [1, 2])
def test_keyed_substeps_inactive_proxy_iterations(vk, iterations, monkeypatch):
    """Later substeps keyed by the previous substep's apply, with inactive and
    proxy particles (the pass-through keys them too) and with two iterations
    (only the last one keys the next substep and resets the cell rows): a pile
    dropped onto an icosphere, against the oracle and against every substep
    keyed by k_integrate."""
    rng = np.random.default_rng(31)
    r = 0.01
    q = scenes.lattice(20, 20, 6, 2 * r, (-0.19, -0.19, 1.01), 0.1 * r, rng)
    flags = np.ones(q.shape[0], np.int32)
    flags[::7] = 0          # inactive
    flags[3::11] = 1 | 2    # active proxy
    verts, tris = scenes.icosphere(3)
    sc = scene_from(q, r, meshes=[(verts, tris)], flags=flags)
    g, o = make_pair(vk, sc)
    monkeypatch.setenv("VKXPBD_FUSE_KEYS", "0")
    h = scenes.build_solver(vk, sc)
    monkeypatch.delenv("VKXPBD_FUSE_KEYS")
    for f in range(8):
        for s in (g, o, h):
            s.step(1.0 / 60.0, 8, iterations)
        assert g.last_contact_count == o.last_contact_count == h.last_contact_count
        if f % 2 == 1:
            compare_state(g, o, exact=True, what=f"keyed, {iterations} iteration(s), frame {f + 1}")
    np.testing.assert_array_equal(g.get_positions(), h.get_positions())
    np.testing.assert_array_equal(g.get_velocities(), h.get_velocities())
    compare_grid(g, o)
    assert_contacts_equal(gpu_contacts(g), o.get_contacts())
    off = flags == 0
    np.testing.assert_array_equal(g.get_positions()[off], q[off])
    assert g.last_contact_count > 0


def test_carried_keys_follow_the_substep_length(vk):
    """A frame's first substep is keyed by the last frame's last substep only
    when the substep length is the same: frames of changing dt and substep
    count, with a mesh and without, against the oracle every frame."""
    rng = np.random.default_rng(41)
    r = 0.01
    q = scenes.lattice(16, 16, 5, 2 * r, (-0.15, -0.15, 1.01), 0.1 * r, rng)
    verts, tris = scenes.icosphere(3)
    for meshes in ([(verts, tris)], []):
        sc = scene_from(q if meshes else q - np.float32([0.0, 0.0, 0.99]), r, meshes=meshes)
        g, o = make_pair(vk, sc)
        for f, (dt, k) in enumerate([(1 / 60, 8), (1 / 60, 8), (1 / 120, 8), (1 / 60, 4), (1 / 60, 4), (1 / 60, 8)]):
            g.step(dt, k, 1)
            o.step(dt, k, 1)
            assert g.last_contact_count == o.last_contact_count
            compare_state(g, o, exact=True, what=f"{len(meshes)} mesh(es), frame {f + 1} (dt {dt:.4f}, {k} substeps)")
        assert_contacts_equal(gpu_contacts(g), o.get_contacts())


def test_update_mesh_moving_collider(vk):
    rng = np.random.default_rng(9)
    r = 0.01
    q = scenes.lattice(16, 16, 4, 2 * r, (-0.15, -0.15, 1.02), 0.1 * r, rng)
    verts, tris = scenes.icosphere(2)
    sc = scene_from(q, r, meshes=[(verts, tris)])
    g, o = make_pair(vk, sc)
    for f in range(12):
        nv = (verts + np.float32([0.0, 0.0, 0.01 * f])).astype(np.float32)
        vel = np.tile(np.float32([0.0, 0.0, 0.6]), (verts.shape[0], 1))
        g.update_mesh(0, nv.reshape(-1), vel.reshape(-1))
        o.update_mesh(0, nv, vel)
        g.step(1.0 / 60.0, 8, 1)
        o.step(1.0 / 60.0, 8, 1)
    lo_g, lo_o = g.get_mesh_bvh_lowers(0), o.get_mesh_bvh(0)[0]
    np.testing.assert_array_equal(lo_g.view(np.uint32), lo_o.view(np.uint32))
    assert_contacts_equal(gpu_contacts(g), o.get_contacts())
    compare_state(g, o, exact=True, what="moving collider")


@pytest.mark.parametrize("nz,planes", [(64, None), (96, None), (96, "0")])
def test_large_grid_properties(vk, nz, planes, monkeypatch):
    """1,048,576 and 1,572,864 particles over the icosphere (512 and 768 radix
    tiles: the bin scan takes one and two steps), with wrap planes and with
    plain cell keys (whose getters return the device's own order):
    size-independent properties of the sort and cell tables after one
    substep with no iterations."""
    sc = scenes.pile_1m_icosphere(nz=nz)
    if planes is not None:
        monkeypatch.setenv("VKXPBD_WRAP_PLANES", planes)
    g = scenes.build_solver(vk, sc)
    assert g.wrap_planes == (planes is None)
    g.step(1.0 / 60.0, 1, 0)   # integrate + grid + sort only
    x = g.get_positions()
    keys, ids = g.get_sorted_cells(), g.get_sorted_point_ids()
    n = x.shape[0]
    assert np.all(keys[1:] >= keys[:-1])
    assert np.array_equal(np.sort(ids), np.arange(n, dtype=np.uint32))
    same = keys[1:] == keys[:-1]
    assert np.all(ids[1:][same] > ids[:-1][same])          # stable: ids ascend inside a cell
    inv = np.float32(1.0) / np.float32(2 * sc["r"])
    c = np.floor(x * inv).astype(np.int64) + (1 << 20)
    c = np.maximum(c, 0) % 128
    cell = (c[:, 2] * 16384 + c[:, 1] * 128 + c[:, 0]).astype(np.uint32)
    np.testing.assert_array_equal(keys, cell[ids])
    st, en = g.get_cell_starts(), g.get_cell_ends()
    assert int(np.sum(en - st)) == n
    u = np.unique(keys)
    np.testing.assert_array_equal(keys[st[u]], u)


def aliased_piles(r=0.01, side=8, layers=4, seed=11):
    """Copies of one small pile placed whole wraps of the 128-cell hash grid
    apart (128 cells along x, y, z and combinations; 512 cells along x, which
    lands in the same x wrap plane; -256 cells along y), so that they share
    hash cells without ever touching, with random velocities."""
    rng = np.random.default_rng(seed)
    cell = 2 * r
    base = scenes.lattice(side, side, layers, 2 * r, (-(side - 1) * r, -(side - 1) * r, 1.2 * r), 0.1 * r, rng)
    shifts = [(0, 0, 0), (128, 0, 0), (0, 128, 0), (0, 0, 128), (128, 128, 0), (-128, 128, 128), (512, 0, 0),
              (0, -256, 0)]
    q = np.concatenate([base + np.float32(np.array(sh, np.float64) * cell) for sh in shifts]).astype(np.float32)
    qd = rng.normal(0.0, 0.2, size=q.shape).astype(np.float32)
    return scene_from(q, r, qd)


def test_wrap_planes_aliased_piles(vk, monkeypatch):
    """Piles that share hash cells across wraps of the grid: the wrap-plane
    keys give the oracle's bits, the same state as plain cell keys, and the
    reference's sorted order and cell tables from the getters."""
    sc = aliased_piles()
    monkeypatch.setenv("VKXPBD_WRAP_PLANES", "1")
    g, o = make_pair(vk, sc)
    monkeypatch.delenv("VKXPBD_WRAP_PLANES")
    h = scenes.build_solver(vk, sc)
    assert g.wrap_planes and not h.wrap_planes
    for f in range(6):
        for s in (g, o, h):
            s.step(1.0 / 60.0, 8, 1)
        compare_state(g, o, exact=True, what=f"aliased piles frame {f + 1}")
        np.testing.assert_array_equal(g.get_positions(), h.get_positions())
        np.testing.assert_array_equal(g.get_velocities(), h.get_velocities())
        assert g.last_contact_count == o.last_contact_count == h.last_contact_count
    compare_grid(g, o)
    np.testing.assert_array_equal(g.get_sorted_point_ids(), h.get_sorted_point_ids())
    np.testing.assert_array_equal(g.get_cell_starts(), h.get_cell_starts())


def test_wrap_planes_follow_the_radii(vk):
    """set_particles with a radius that could reach across a wrap turns the
    planes off, a later small radius turns them on again; every frame keeps
    the oracle's bits."""
    sc = aliased_piles(side=6, layers=3)
    g, o = make_pair(vk, sc)
    p = sc["particles"]
    assert not g.wrap_planes
    g.set_wrap_planes(True)
    assert g.wrap_planes
    big = p["radii"].copy()
    big[5] = 0.9   # 3 * 0.9 + 0.01 >= 126 * 0.02
    for f, radii in enumerate([None, big, None, p["radii"]]):
        if radii is not None:
            x, v = g.get_positions(), g.get_velocities()
            for s in (g, o):
                s.set_particles(x, v, p["inv_mass"], radii, p["flags"])
            assert g.wrap_planes == (radii is p["radii"])
        for s in (g, o):
            s.step(1.0 / 60.0, 4, 1)
        compare_state(g, o, exact=True, what=f"radii phase {f} planes={g.wrap_planes}")
        compare_grid(g, o)


def random_quats(rng, n):
    """Uniform random unit quaternions (Shoemake), xyzw."""
    u1, u2, u3 = rng.uniform(size=(3, n))
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], 1)
    return q.astype(np.float32)


def test_boxes_and_particles(vk):
    """Rigid boxes (integrate_bodies, box-plane rigid contacts,
    solve_body_contact_positions, apply_body_deltas) with particles colliding
    against them.  Body deltas are float atomics summed in arbitrary order
    (several corners / particles per body), so the bar is 1e-5 after a frame,
    with the later drift reported."""
    rng = np.random.default_rng(12)
    r = 0.005
    nb = 27
    pos = np.stack(np.meshgrid(np.arange(3), np.arange(3), np.arange(3), indexing="ij"), -1).reshape(-1, 3)
    pos = (pos * [0.25, 0.25, 0.2] + [-0.25, -0.25, 0.1]).astype(np.float32)
    quats = random_quats(rng, nb)
    q = scenes.lattice(24, 24, 6, 2 * r, (-0.115, -0.115, 0.7), 0.1 * r, rng)
    sc = scene_from(q, r)
    cfg = scenes.solver_config(vk, r)
    g = vk.XpbdSolver()
    o = O.OracleSolver()
    g.init(cfg, q.shape[0])
    o.init(cfg, q.shape[0])
    for k in range(nb):
        g.add_box_body(pos[k], quats[k], [0.05, 0.05, 0.05], 1000.0)
        o.add_box_body(pos[k], quats[k], [0.05, 0.05, 0.05], 1000.0)
    g.finalize_meshes()
    o.finalize_meshes()
    p = sc["particles"]
    g.set_particles(p["q"], p["qd"], p["inv_mass"], p["radii"], p["flags"])
    o.set_particles(p["q"], p["qd"], p["inv_mass"], p["radii"], p["flags"])
    g.step(1.0 / 60.0, 8, 1)
    o.step(1.0 / 60.0, 8, 1)
    db = float(np.max(np.abs(g.get_body_q().astype(np.float64) - o.get_body_q())))
    log(f"  boxes frame 1: max|dq_body|={db:.3e}")
    assert db <= TOL_M
    compare_state(g, o, exact=False, what="boxes frame 1 (particles)")
    drift = []
    for f in range(40):
        g.step(1.0 / 60.0, 8, 1)
        o.step(1.0 / 60.0, 8, 1)
        if f % 10 == 9:
            drift.append(float(np.max(np.abs(g.get_body_q().astype(np.float64) - o.get_body_q()))))
    log(f"  body pose drift every 10 frames: {drift}")
    gq = g.get_body_q()
    assert np.all(np.isfinite(gq)) and np.all(gq[:, 2] > 0.0)        # boxes rest above the plane
    np.testing.assert_allclose(np.linalg.norm(gq[:, 3:], axis=1), 1.0, atol=1e-5)


def test_512_boxes_and_262144_particles(vk):
    """The rigid-body configuration at full size: 512 boxes with random
    orientations falling onto the plane and 262,144 particles falling onto
    them.  After the first frame the state is within 1e-5 m of the oracle
    (body deltas are float atomics summed in arbitrary order); later frames'
    drift is reported, and the boxes stay normalised and above the plane."""
    import time
    sc = scenes.boxes_and_pile()
    g, o = make_pair(vk, sc)
    t0 = time.perf_counter()
    g.step(1.0 / 60.0, 8, 1)
    t1 = time.perf_counter()
    o.step(1.0 / 60.0, 8, 1)
    log(f"boxes and pile frame 1: gpu {1e3 * (t1 - t0):.1f} ms (first frame, includes capture); oracle {time.perf_counter() - t1:.1f} s")
    dx = compare_state(g, o, exact=False, what="boxes and pile frame 1")
    db = float(np.max(np.abs(g.get_body_q().astype(np.float64) - o.get_body_q())))
    log(f"  bodies frame 1: max|dq_body|={db:.3e}")
    assert db <= 1e-5 and dx <= 1e-5
    for f in range(2, 6):
        g.step(1.0 / 60.0, 8, 1)
        o.step(1.0 / 60.0, 8, 1)
        gp, op = g.get_positions().astype(np.float64), o.get_positions()
        db = float(np.max(np.abs(g.get_body_q().astype(np.float64) - o.get_body_q())))
        log(f"  frame {f}: particle drift {float(np.max(np.abs(gp - op))):.3e} m, body drift {db:.3e}")
    bq = g.get_body_q()
    assert np.allclose(np.linalg.norm(bq[:, 3:], axis=1), 1.0, atol=1e-5)
    assert np.all(bq[:, 2] > 0.0)
    t0 = time.perf_counter()
    for _ in range(10):
        g.step(1.0 / 60.0, 8, 1)
    log(f"  gpu {1e3 * (time.perf_counter() - t0) / 10:.2f} ms per frame (frames 7-16)")


def test_graph_replay_equals_eager(vk):
    """Captured frames (hipGraph replay) give the same bits as eager launches,
    across an update_mesh (which drops the captured frames) and a change of
    the step parameters."""
    sc = scenes.pile_4096()
    v, t = scenes.icosphere(2, 0.1, (0.0, 0.0, 0.1))
    sc["meshes"] = [(v, t)]
    a = scenes.build_solver(vk, sc)
    b = scenes.build_solver(vk, sc)
    b.set_graph_mode(False)
    for f in range(12):
        if f == 6:
            v2 = v + np.float32([0.01, 0.0, 0.0])
            a.update_mesh(0, v2.reshape(-1))
            b.update_mesh(0, v2.reshape(-1))
        it = 1 if f < 9 else 2
        a.step(1.0 / 60.0, 8, it)
        b.step(1.0 / 60.0, 8, it)
        np.testing.assert_array_equal(a.get_positions(), b.get_positions())
        np.testing.assert_array_equal(a.get_velocities(), b.get_velocities())
        assert a.last_contact_count == b.last_contact_count
    log(f"graph replay == eager over 12 frames, contacts {a.last_contact_count}")


@pytest.mark.parametrize("num_scenes,lo,hi,frames", [(7, 64, 3000, 20), (300, 16, 64, 4)])
def test_batched_scenes_match_sequential(vk, num_scenes, lo, hi, frames):
    """A batched solver over uneven scenes equals one plain solver per scene,
    bit for bit (each scene keeps its own grid, neighbour order and plane).
    7 scenes sort 24-bit keys in three 8-bit digits, 300 scenes 30-bit keys
    in four."""
    sc = scenes.batch_of_piles(seed=9, num_scenes=num_scenes, lo=lo, hi=hi)
    sc["counts"][3] = 1   # a single-particle scene
    off = np.concatenate([[0], np.cumsum(sc["counts"])])
    sc["particles"] = {k: np.concatenate([scenes.scene_slice(sc, j)["particles"][k] for j in range(num_scenes)])
                       for k in sc["particles"]}
    log(f"batched scenes: {num_scenes}, {int(off[-1])} particles")
    bs = scenes.build_batch_solver(vk, sc)
    assert bs.scene_count == num_scenes
    singles = [scenes.build_solver(vk, scenes.scene_slice(sc, k)) for k in range(num_scenes)]
    for f in range(frames):
        bs.step(1.0 / 60.0, 8, 1)
        tot = 0
        for s in singles:
            s.step(1.0 / 60.0, 8, 1)
            tot += s.last_contact_count
        gp = bs.get_positions()
        gv = bs.get_velocities()
        for k, s in enumerate(singles):
            np.testing.assert_array_equal(gp[off[k]:off[k + 1]], s.get_positions(), err_msg=f"scene {k} frame {f}")
            np.testing.assert_array_equal(gv[off[k]:off[k + 1]], s.get_velocities(), err_msg=f"scene {k} frame {f}")
        assert bs.last_contact_count == tot
    # the cell ranges: a row table and group bounds, not a cell table per scene
    assert bs.cell_table_bytes < 64 * 1024 * 1024
    # scenes overlap in space but never interact: keys carry the scene index
    keys = bs.get_sorted_cells()
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)
    assert int(keys.max()) >> 21 == num_scenes - 1


@pytest.mark.parametrize("env", [{"VKXPBD_SEG_SORT": "0"},
                                 {"VKXPBD_SEG_SORT": "0", "VKXPBD_SORT_CHUNKED": "1"},
                                 {"VKXPBD_BATCH_GATHER_IN_ROWS": "0"}],
                         ids=["whole_batch_radix", "whole_batch_chunked_radix", "gather_in_cell_ranges"])
def test_batched_sort_variants(vk, env, monkeypatch):
    """A batched solver whose radix passes sort the whole batch on
    scene-tagged keys (plain and chunked) instead of each scene in its own
    range, or whose state move runs in the cell-range pass instead of the
    row-table build, equals the default batched solver and one plain solver
    per scene, bit for bit."""
    num_scenes, frames = 9, 6
    sc = scenes.batch_of_piles(seed=17, num_scenes=num_scenes, lo=64, hi=6000)
    sc["counts"][2] = 1
    off = np.concatenate([[0], np.cumsum(sc["counts"])])
    sc["particles"] = {k: np.concatenate([scenes.scene_slice(sc, j)["particles"][k] for j in range(num_scenes)])
                       for k in sc["particles"]}
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    bs = scenes.build_batch_solver(vk, sc)
    for k in env:
        monkeypatch.delenv(k)
    d = scenes.build_batch_solver(vk, sc)   # the defaults
    singles = [scenes.build_solver(vk, scenes.scene_slice(sc, k)) for k in range(num_scenes)]
    log(f"batched variant {env}: {int(off[-1])} particles")
    for f in range(frames):
        bs.step(1.0 / 60.0, 8, 1)
        d.step(1.0 / 60.0, 8, 1)
        for s in singles:
            s.step(1.0 / 60.0, 8, 1)
        gp, gv = bs.get_positions(), bs.get_velocities()
        np.testing.assert_array_equal(gp, d.get_positions(), err_msg=f"frame {f}")
        np.testing.assert_array_equal(gv, d.get_velocities(), err_msg=f"frame {f}")
        assert bs.last_contact_count == d.last_contact_count == sum(s.last_contact_count for s in singles)
        for k, s in enumerate(singles):
            np.testing.assert_array_equal(gp[off[k]:off[k + 1]], s.get_positions(), err_msg=f"scene {k} frame {f}")
            np.testing.assert_array_equal(gv[off[k]:off[k + 1]], s.get_velocities(), err_msg=f"scene {k} frame {f}")
    keys = bs.get_sorted_cells()
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)


def test_2048_scenes_use_every_scene_bit(vk):
    """The largest batch (2,048 scenes, keys up to 2^32 - 1): the first, a
    middle and the last scenes equal plain solvers bit for bit, and the top
    scene's cells sort last."""
    num_scenes = 2048
    sc = scenes.batch_of_piles(seed=17, num_scenes=num_scenes, lo=8, hi=24)
    off = np.concatenate([[0], np.cumsum(sc["counts"])])
    bs = scenes.build_batch_solver(vk, sc)
    assert bs.scene_count == num_scenes
    picks = [0, 1, 1023, 2046, 2047]
    singles = {k: scenes.build_solver(vk, scenes.scene_slice(sc, k)) for k in picks}
    for f in range(3):
        bs.step(1.0 / 60.0, 8, 1)
        for s in singles.values():
            s.step(1.0 / 60.0, 8, 1)
        gp, gv = bs.get_positions(), bs.get_velocities()
        for k, s in singles.items():
            np.testing.assert_array_equal(gp[off[k]:off[k + 1]], s.get_positions(), err_msg=f"scene {k} frame {f}")
            np.testing.assert_array_equal(gv[off[k]:off[k + 1]], s.get_velocities(), err_msg=f"scene {k} frame {f}")
    keys = bs.get_sorted_cells()
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)
    assert int(keys.max()) >> 21 == num_scenes - 1
    log(f"2048 scenes, {int(off[-1])} particles, cell tables {bs.cell_table_bytes / 2**20:.1f} MiB")


def test_every_tuning_gives_the_same_bits(vk):
    """Each launch shape autotune may pick (pp block 64/128/256 x radix tile
    1,024/2,048 keys) gives the default's bits: full frames (captured), then
    a step_solve_only over the last contact set."""
    sc = scenes.pile_4096()
    ref = scenes.build_solver(vk, sc)
    for _ in range(6):
        ref.step(1.0 / 60.0, 8, 1)
    ref.step_solve_only(1.0 / 60.0, 2, 1)
    want_x, want_v = ref.get_positions(), ref.get_velocities()
    for block in (64, 128, 256):
        for tile in (1024, 2048):
            s = scenes.build_solver(vk, sc)
            s.set_tuning(block, tile)
            assert (s.pp_block, s.radix_tile) == (block, tile)
            for _ in range(6):
                s.step(1.0 / 60.0, 8, 1)
            s.step_solve_only(1.0 / 60.0, 2, 1)
            log(f"  block {block} tile {tile}: max|dx| {float(np.max(np.abs(s.get_positions() - want_x))):.3e}")
            np.testing.assert_array_equal(s.get_positions(), want_x, err_msg=f"block {block} tile {tile}")
            np.testing.assert_array_equal(s.get_velocities(), want_v, err_msg=f"block {block} tile {tile}")
    with pytest.raises(Exception, match="pp_block"):
        ref.set_tuning(32, 1024)
    with pytest.raises(Exception, match="radix_tile"):
        ref.set_tuning(64, 4096)


def test_autotune_keeps_state(vk):
    """autotune runs trial frames and puts everything back: right after it,
    the contacts, sorted cells, deltas and a step_solve_only over the last
    contact set equal an untuned solver's, and so do later full steps."""
    sc = scenes.pile_4096()
    a = scenes.build_solver(vk, sc)
    b = scenes.build_solver(vk, sc)
    for s in (a, b):
        for _ in range(10):
            s.step(1.0 / 60.0, 8, 1)
    blk = a.autotune(2, 1.0 / 60.0, 8, 1)
    log(f"autotune picked pp block {blk}")
    assert blk in (64, 128, 256)
    assert a.last_contact_count == b.last_contact_count > 0
    assert_contacts_equal(gpu_contacts(a), gpu_contacts(b))
    np.testing.assert_array_equal(a.get_sorted_cells(), b.get_sorted_cells())
    np.testing.assert_array_equal(a.get_sorted_point_ids(), b.get_sorted_point_ids())
    np.testing.assert_array_equal(a.get_cell_starts(), b.get_cell_starts())
    np.testing.assert_array_equal(a.get_deltas(), b.get_deltas())
    a.step_solve_only(1.0 / 60.0, 2, 1)
    b.step_solve_only(1.0 / 60.0, 2, 1)
    np.testing.assert_array_equal(a.get_positions(), b.get_positions())
    for _ in range(3):
        a.step(1.0 / 60.0, 8, 1)
        b.step(1.0 / 60.0, 8, 1)
    np.testing.assert_array_equal(a.get_positions(), b.get_positions())
    np.testing.assert_array_equal(a.get_velocities(), b.get_velocities())


def test_radius_change_after_captured_frame(vk):
    """A captured frame bakes the particle-pair reject bound (the largest
    radius) in: set_particles with larger radii must drop it, so the next
    frame keeps the oracle's bits and equals eager launches."""
    sc = scenes.pile_4096()
    g, o = make_pair(vk, sc)
    e = scenes.build_solver(vk, sc)
    e.set_graph_mode(False)
    for s in (g, o, e):
        s.step(1.0 / 60.0, 8, 1)
    p = sc["particles"]
    big = (p["radii"] * np.float32(1.5)).astype(np.float32)
    x, v = g.get_positions(), g.get_velocities()
    for s in (g, o, e):
        s.set_particles(x, v, p["inv_mass"], big, p["flags"])
    for f in range(3):
        for s in (g, o, e):
            s.step(1.0 / 60.0, 8, 1)
        compare_state(g, o, exact=True, what=f"radii x1.5, frame {f + 1}")
        np.testing.assert_array_equal(g.get_positions(), e.get_positions())
        assert g.last_contact_count == o.last_contact_count
