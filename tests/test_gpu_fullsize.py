"""The benchmark workloads at their full sizes on the GPU, against the oracle
and against the plain solver.

* The 1,048,576-particle pile over the 81,920-triangle icosphere: the mesh
  narrowphase at this mesh resolution (distance-bound grid, start nodes,
  four-lane walk, parity grid) against the oracle's tree search and
  all-triangle ray parity (generate_mesh_contacts.comp:181-414), on a shell
  of particles around the sphere and over the first ten frames of the pile.
* The 256 batched piles (3,813,381 particles): the batched solver against
  256 plain solvers and, for four of its scenes, against the oracle.

The oracle's loops run on all host threads (results do not depend on the
thread count); each test prints its progress per frame."""
import time

import numpy as np
import pytest

import oracle_lib as O
from gpu_util import assert_contacts_equal, gpu_contacts, log, make_pair
from vkxpbd import scenes

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def compare_state(g, o, what):
    gp, op = g.get_positions().astype(np.float64), o.get_positions()
    gv, ov = g.get_velocities().astype(np.float64), o.get_velocities()
    log(f"  {what}: max|dx|={float(np.max(np.abs(gp - op))):.3e} max|dv|={float(np.max(np.abs(gv - ov))):.3e}")
    np.testing.assert_array_equal(gp, op, err_msg=what)
    np.testing.assert_array_equal(gv, ov, err_msg=what)


def test_icosphere81920_shell_collide_only(vk):
    """50,000 particles within 2 cm of the radius-0.5 icosphere of 81,920
    triangles, inside and outside: collide_only's contact set, closest
    points, normals (watertight sign) and interpolated velocities equal the
    oracle's bit for bit."""
    verts, tris = scenes.icosphere(6)
    assert tris.shape[0] == 81920
    rng = np.random.default_rng(61)
    n = 50000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    q = (d * (0.5 + rng.uniform(-0.02, 0.02, size=(n, 1))) + [0.0, 0.0, 0.5]).astype(np.float32)
    r = 0.005
    sc = dict(name="shell", r=r, meshes=[(verts, tris)],
              particles=dict(q=q, qd=np.zeros_like(q), inv_mass=np.full(n, scenes.PARTICLE_INV_MASS, np.float32),
                             radii=np.full(n, r, np.float32), flags=np.ones(n, np.int32)))
    g, o = make_pair(vk, sc)
    g.collide_only()
    t0 = time.perf_counter()
    o.collide_only()
    log(f"shell: contacts gpu={g.last_contact_count} oracle={o.last_contact_count} "
        f"(oracle {time.perf_counter() - t0:.1f} s)")
    assert g.last_contact_count == o.last_contact_count > 30000
    c = gpu_contacts(g)
    assert_contacts_equal(c, o.get_contacts())
    m = c["shapes"] == 1
    inside = np.linalg.norm(q[c["particles"][m]] - [0.0, 0.0, 0.5], axis=1) < 0.499
    sd = np.sum((q[c["particles"][m]] - c["body_positions"][m]) * c["normals"][m], axis=1)
    assert np.all(sd[inside] < 0.0) and np.any(inside)


def test_pile_1m_icosphere_first_ten_frames(vk):
    """Frames 1-10 of the 1,048,576-particle pile over the 81,920-triangle
    icosphere and the plane: the pile reaches the sphere at frame 5, and at
    frames 8 and 10 the contact sets and the whole state equal the oracle's
    bit for bit (contact counts every frame)."""
    sc = scenes.pile_1m_icosphere()
    g, o = make_pair(vk, sc)
    assert g.wrap_planes
    for f in range(1, 11):
        g.step(DT, 8, 1)
        t0 = time.perf_counter()
        o.step(DT, 8, 1)
        log(f"frame {f}: contacts gpu={g.last_contact_count} oracle={o.last_contact_count} "
            f"(oracle {time.perf_counter() - t0:.1f} s)")
        assert g.last_contact_count == o.last_contact_count
        if f in (8, 10):
            c = gpu_contacts(g)
            assert np.count_nonzero(c["shapes"] == 1) > 1000
            assert_contacts_equal(c, o.get_contacts())
            compare_state(g, o, f"frame {f}")


def test_batch_256_piles_full_size(vk):
    """The 256-scene batch (counts log-uniform on [1024, 65536], 3,813,381
    particles): the batched solver's cell ranges take under 0.5 GiB (a row
    table of 8 bytes per grid row and group bounds of at most 36 bytes per
    particle, not a cell table per scene); after autotuning, two frames of it equal
    256 plain solvers bit for bit, and the smallest, the largest and two
    seeded scenes equal the oracle bit for bit after the second frame (plane
    and particle contacts)."""
    sc = scenes.batch_of_piles()
    counts = np.asarray(sc["counts"])
    off = np.concatenate([[0], np.cumsum(counts)])
    log(f"batch: {counts.size} scenes, {int(off[-1])} particles")
    assert counts.size == 256 and off[-1] > 3_000_000
    bs = scenes.build_batch_solver(vk, sc)
    log(f"batched cell-range tables: {bs.cell_table_bytes / 2**20:.1f} MiB")
    assert bs.cell_table_bytes < 512 * 1024 * 1024
    block = bs.autotune(1, DT, 8, 1)
    log(f"autotune: particle-pair block {block}, radix tile {bs.radix_tile} keys")
    singles = [scenes.build_solver(vk, scenes.scene_slice(sc, k)) for k in range(counts.size)]
    picks = sorted({int(np.argmin(counts)), int(np.argmax(counts))} |
                   set(int(k) for k in np.random.default_rng(3).choice(counts.size, 2, replace=False)))
    oracles = {}
    for k in picks:
        p = scenes.scene_slice(sc, k)["particles"]
        o = O.OracleSolver()
        o.init(scenes.solver_config(vk, sc["r"]), p["q"].shape[0])
        o.finalize_meshes()
        o.set_particles(p["q"], p["qd"], p["inv_mass"], p["radii"], p["flags"])
        oracles[k] = o
    for f in range(1, 3):
        bs.step(DT, 8, 1)
        tot = 0
        for s in singles:
            s.step(DT, 8, 1)
            tot += s.last_contact_count
        assert bs.last_contact_count == tot
        gp, gv = bs.get_positions(), bs.get_velocities()
        for k, s in enumerate(singles):
            np.testing.assert_array_equal(gp[off[k]:off[k + 1]], s.get_positions(), err_msg=f"scene {k} frame {f}")
            np.testing.assert_array_equal(gv[off[k]:off[k + 1]], s.get_velocities(), err_msg=f"scene {k} frame {f}")
        log(f"frame {f}: batched == 256 plain solvers, contacts {tot}")
        for k, o in oracles.items():
            o.step(DT, 8, 1)
        if f == 2:   # the piles reach the plane during frame 2
            assert tot > 0
            for k, o in oracles.items():
                np.testing.assert_array_equal(gp[off[k]:off[k + 1]].astype(np.float64), o.get_positions(),
                                              err_msg=f"scene {k} vs oracle")
                np.testing.assert_array_equal(gv[off[k]:off[k + 1]].astype(np.float64), o.get_velocities(),
                                              err_msg=f"scene {k} vs oracle")
                log(f"  scene {k} ({counts[k]} particles) == oracle")
