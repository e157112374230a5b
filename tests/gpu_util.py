"""Shared helpers for the GPU parity tests: run a scene on the HIP solver and
on the oracle with identical inputs, and canonicalise contact lists."""
import sys

import numpy as np

import oracle_lib as O
from vkxpbd import scenes


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_pair(vk, scene, max_contacts=0, fp64=False, **cfg_overrides):
    cfg = scenes.solver_config(vk, scene["r"], **cfg_overrides)
    g = scenes.build_solver(vk, scene, max_contacts=max_contacts, **cfg_overrides)
    o = O.OracleSolver(fp64=fp64)
    p = scene["particles"]
    o.init(cfg, p["q"].shape[0], max_contacts)
    for verts, tris in scene["meshes"]:
        o.add_mesh(verts, tris)
    for pos, quat, half, density in scene.get("boxes", []):
        o.add_box_body(pos, quat, half, density)
    o.finalize_meshes()
    o.set_particles(p["q"], p["qd"], p["inv_mass"], p["radii"], p["flags"])
    return g, o


def gpu_contacts(s):
    return dict(particles=s.get_contact_particles(), shapes=s.get_contact_shapes(),
                body_positions=s.get_contact_body_positions().astype(np.float64),
                body_velocities=s.get_contact_body_velocities().astype(np.float64),
                normals=s.get_contact_normals().astype(np.float64))


def canonical(c):
    """Contacts sorted by (particle, shape): unique, since there is at most one
    contact per particle and shape; the reference's slot order is arbitrary."""
    order = np.lexsort((c["shapes"], c["particles"]))
    return {k: np.asarray(v)[order] for k, v in c.items()}


def assert_contacts_equal(a, b):
    a, b = canonical(a), canonical(b)
    np.testing.assert_array_equal(a["particles"], b["particles"])
    np.testing.assert_array_equal(a["shapes"], b["shapes"])
    for k in ("body_positions", "body_velocities", "normals"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def scene_from(q, r, qd=None, meshes=(), inv_mass=scenes.PARTICLE_INV_MASS, flags=None):
    n = q.shape[0]
    return dict(name="custom", r=r, meshes=list(meshes),
                particles=dict(q=np.ascontiguousarray(q, np.float32),
                               qd=np.zeros_like(q, np.float32) if qd is None else np.ascontiguousarray(qd, np.float32),
                               inv_mass=np.full(n, inv_mass, np.float32), radii=np.full(n, r, np.float32),
                               flags=np.ones(n, np.int32) if flags is None else np.asarray(flags, np.int32)))
