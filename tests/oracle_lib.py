"""ctypes access to the CPU restatement (oracle/liboracle.so) and, where it
was built, to the reference's own LBVH code (oracle/_ref/libref_lbvh.so).

Test infrastructure only: the oracle is the checker, never the thing measured.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_lbvh.so")
NUM_CELLS = 128 ** 3

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)


class OrcConfig(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "particle_v_max", "soft_contact_mu", "shape_material_mu", "soft_contact_relaxation", "particle_adhesion",
        "particle_mu", "particle_cohesion", "particle_max_radius", "search_radius", "soft_contact_margin",
        "soft_contact_restitution", "ground_plane_altitude")] + [
        ("gravity", C.c_float * 3), ("enable_restitution", C.c_int32), ("has_ground_plane", C.c_int32)]


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/liboracle.so not built (run `make -C oracle`)")
        L = C.CDLL(LIB)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_last_error.restype = C.c_char_p
        L.orc_last_error.argtypes = [C.c_void_p]
        L.orc_init.argtypes = [C.c_void_p, C.POINTER(OrcConfig), C.c_int64, C.c_int64]
        L.orc_add_mesh.argtypes = [C.c_void_p, _fp, C.c_int64, _up, C.c_int64, C.c_float, C.c_int]
        L.orc_finalize_meshes.argtypes = [C.c_void_p]
        L.orc_update_mesh.argtypes = [C.c_void_p, C.c_int64, _fp, C.c_int64, _fp]
        L.orc_set_particles.argtypes = [C.c_void_p, _fp, _fp, _fp, _fp, _ip, C.c_int64]
        L.orc_step.argtypes = [C.c_void_p, C.c_float, C.c_int, C.c_int]
        L.orc_step_solve_only.argtypes = [C.c_void_p, C.c_float, C.c_int, C.c_int]
        L.orc_collide_only.argtypes = [C.c_void_p]
        for f in ("orc_get_positions", "orc_get_velocities", "orc_get_deltas"):
            getattr(L, f).argtypes = [C.c_void_p, _dp]
        L.orc_contact_count.restype = C.c_int32
        L.orc_contact_count.argtypes = [C.c_void_p]
        L.orc_get_contacts.argtypes = [C.c_void_p, _ip, _ip, _dp, _dp, _dp]
        L.orc_set_contact_data.argtypes = [C.c_void_p, _ip, _ip, _fp, _fp, C.c_int32]
        L.orc_get_sorted.argtypes = [C.c_void_p, _up, _up]
        L.orc_get_cell_tables.argtypes = [C.c_void_p, _ip, _ip]
        L.orc_mesh_node_count.restype = C.c_int64
        L.orc_mesh_node_count.argtypes = [C.c_void_p, C.c_int64]
        L.orc_get_mesh_bvh.argtypes = [C.c_void_p, C.c_int64, _fp, _fp, _up, _up]
        L.orc_mesh_watertight.argtypes = [C.c_void_p, C.c_int64]
        L.orc_add_box_body.argtypes = [C.c_void_p, _fp, _fp, _fp, C.c_float, C.c_int32]
        L.orc_set_body_params.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int]
        L.orc_get_bodies.restype = C.c_int64
        L.orc_get_bodies.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_set_body_velocities.argtypes = [C.c_void_p, _fp]
        L.orc_lbvh_build.argtypes = [_fp, _up, C.c_uint32, _fp, _fp, _up, _up]
        L.orc_ray_tri_woop_f.argtypes = [_fp] * 5
        L.orc_closest_point_triangle_f.argtypes = [_fp] * 6
        L.orc_rollout.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, C.c_int, C.c_int]
        L.orc_rollout_backward.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.orc_ray_hit_lists.restype = C.c_int64
        L.orc_ray_hit_lists.argtypes = [_fp, _up, C.c_int64, _fp, C.c_int64, _ip, _up, C.c_int64]
        _lib = L
    return _lib


def config_from(cfg) -> OrcConfig:
    """OrcConfig from a vkxpbd.SolverConfig (or any object with the same fields)."""
    o = OrcConfig()
    for name, _ in OrcConfig._fields_:
        if name == "gravity":
            o.gravity[:] = list(cfg.gravity)
        else:
            setattr(o, name, type(getattr(o, name))(getattr(cfg, name)))
    return o


class OracleSolver:
    """The restated XpbdSolver on the CPU; fp64=True gives the FP64 host step."""

    def __init__(self, fp64=False):
        self.L = lib()
        self.h = self.L.orc_create(1 if fp64 else 0)
        self.n = 0

    def __del__(self):
        try:
            self.L.orc_destroy(self.h)
        except Exception:
            pass

    def _ck(self, rc):
        if rc:
            raise RuntimeError(self.L.orc_last_error(self.h).decode())

    def init(self, cfg, n, max_contacts=0):
        self.cfg = config_from(cfg)
        self.n = n
        self._ck(self.L.orc_init(self.h, C.byref(self.cfg), n, max_contacts))

    def add_mesh(self, verts, tris, shape_margin=0.0, is_watertight=None):
        v = np.ascontiguousarray(verts, np.float32).reshape(-1)
        t = np.ascontiguousarray(tris, np.uint32).reshape(-1)
        wt = -1 if is_watertight is None else int(bool(is_watertight))
        self._ck(self.L.orc_add_mesh(self.h, _ptr(v, _fp), v.size // 3, _ptr(t, _up), t.size // 3, shape_margin, wt))

    def finalize_meshes(self):
        self._ck(self.L.orc_finalize_meshes(self.h))

    def rollout(self, v0, relax, sdt, substeps, keep_tape=True):
        """Taped rollout of `substeps` one-iteration substeps from the current
        positions with initial velocities v0 ((N,3), None keeps the current)."""
        v = None if v0 is None else np.ascontiguousarray(v0, np.float64).reshape(-1)
        self._ck(self.L.orc_rollout(self.h, _ptr(v, _dp) if v is not None else None, float(relax), float(sdt),
                                    int(substeps), 1 if keep_tape else 0))

    def rollout_backward(self, gx, gv=None):
        """FP64 adjoint of the last taped rollout: (gx0, gv0, grelax)."""
        gx = np.ascontiguousarray(gx, np.float64).reshape(-1)
        gvv = None if gv is None else np.ascontiguousarray(gv, np.float64).reshape(-1)
        gx0 = np.zeros_like(gx)
        gv0 = np.zeros_like(gx)
        gr = np.zeros(1, np.float64)
        self._ck(self.L.orc_rollout_backward(self.h, _ptr(gx, _dp), _ptr(gvv, _dp) if gvv is not None else None,
                                             _ptr(gx0, _dp), _ptr(gv0, _dp), _ptr(gr, _dp)))
        return gx0.reshape(-1, 3), gv0.reshape(-1, 3), float(gr[0])

    def update_mesh(self, idx, verts, vels=None):
        v = np.ascontiguousarray(verts, np.float32).reshape(-1)
        w = None if vels is None else np.ascontiguousarray(vels, np.float32).reshape(-1)
        self._ck(self.L.orc_update_mesh(self.h, idx, _ptr(v, _fp), v.size // 3, _ptr(w, _fp)))

    def add_box_body(self, pos, quat, half_extents, density=1000.0, flags=0):
        a = [np.ascontiguousarray(x, np.float32) for x in (pos, quat, half_extents)]
        self.nbodies = getattr(self, "nbodies", 0) + 1
        self.L.orc_add_box_body(self.h, *[_ptr(x, _fp) for x in a], density, flags)

    def set_body_params(self, angular_damping=0.05, rigid_contact_relaxation=0.8, mu_torsional=0.0, mu_rolling=0.0,
                        use_constraint_weights=True):
        self.L.orc_set_body_params(self.h, angular_damping, rigid_contact_relaxation, mu_torsional, mu_rolling,
                                   int(use_constraint_weights))

    def get_body_q(self):
        nb = getattr(self, "nbodies", 0)
        q = np.zeros((nb, 7), np.float64)
        self.L.orc_get_bodies(self.h, _ptr(q, _dp), None)
        return q

    def get_body_qd(self):
        nb = getattr(self, "nbodies", 0)
        qd = np.zeros((nb, 6), np.float64)
        self.L.orc_get_bodies(self.h, None, _ptr(qd, _dp))
        return qd

    def set_body_velocities(self, qd):
        a = np.ascontiguousarray(qd, np.float32).reshape(-1)
        self.L.orc_set_body_velocities(self.h, _ptr(a, _fp))

    def set_particles(self, q, qd=None, inv_mass=None, radii=None, flags=None):
        q = np.ascontiguousarray(q, np.float32).reshape(-1)
        conv = lambda a, t: None if a is None else np.ascontiguousarray(a, t).reshape(-1)
        qd, w, r, f = conv(qd, np.float32), conv(inv_mass, np.float32), conv(radii, np.float32), conv(flags, np.int32)
        self._keep = (q, qd, w, r, f)
        self._ck(self.L.orc_set_particles(self.h, _ptr(q, _fp), _ptr(qd, _fp), _ptr(w, _fp), _ptr(r, _fp),
                                          _ptr(f, _ip), q.size // 3))

    def step(self, dt=1.0 / 60.0, substeps=1, iterations=2):
        self._ck(self.L.orc_step(self.h, dt, substeps, iterations))

    def step_solve_only(self, dt, substeps, iterations):
        self._ck(self.L.orc_step_solve_only(self.h, dt, substeps, iterations))

    def collide_only(self):
        self._ck(self.L.orc_collide_only(self.h))

    def _v3(self, fn):
        out = np.zeros((self.n, 3), np.float64)
        fn(self.h, _ptr(out, _dp))
        return out

    def get_positions(self):
        return self._v3(self.L.orc_get_positions)

    def get_velocities(self):
        return self._v3(self.L.orc_get_velocities)

    def get_deltas(self):
        return self._v3(self.L.orc_get_deltas)

    @property
    def last_contact_count(self):
        return int(self.L.orc_contact_count(self.h))

    def get_contacts(self):
        c = self.last_contact_count
        part = np.zeros(c, np.int32)
        shape = np.zeros(c, np.int32)
        bp, bv, nn = (np.zeros((c, 3), np.float64) for _ in range(3))
        if c:
            self.L.orc_get_contacts(self.h, _ptr(part, _ip), _ptr(shape, _ip), _ptr(bp, _dp), _ptr(bv, _dp),
                                    _ptr(nn, _dp))
        return dict(particles=part, shapes=shape, body_positions=bp, body_velocities=bv, normals=nn)

    def set_contact_data(self, particles, shapes, body_positions, normals, count):
        p = np.ascontiguousarray(particles, np.int32)
        s = np.ascontiguousarray(shapes, np.int32)
        b = np.ascontiguousarray(body_positions, np.float32).reshape(-1)
        nn = np.ascontiguousarray(normals, np.float32).reshape(-1)
        self.L.orc_set_contact_data(self.h, _ptr(p, _ip), _ptr(s, _ip), _ptr(b, _fp), _ptr(nn, _fp), count)

    def get_sorted(self):
        k = np.zeros(self.n, np.uint32)
        v = np.zeros(self.n, np.uint32)
        self.L.orc_get_sorted(self.h, _ptr(k, _up), _ptr(v, _up))
        return k, v

    def get_cell_tables(self):
        s = np.zeros(NUM_CELLS, np.int32)
        e = np.zeros(NUM_CELLS, np.int32)
        self.L.orc_get_cell_tables(self.h, _ptr(s, _ip), _ptr(e, _ip))
        return s, e

    def get_mesh_bvh(self, mi):
        nodes = int(self.L.orc_mesh_node_count(self.h, mi))
        lo = np.zeros((nodes, 4), np.float32)
        hi = np.zeros((nodes, 4), np.float32)
        prim = np.zeros(nodes // 2, np.uint32)
        root = np.zeros(1, np.uint32)
        self.L.orc_get_mesh_bvh(self.h, mi, _ptr(lo, _fp), _ptr(hi, _fp), _ptr(prim, _up), _ptr(root, _up))
        return lo, hi, prim, int(root[0])


def lbvh_build(verts, tris):
    """Oracle LBVH restatement: (lowers[2T,4], uppers[2T,4], prim[T], root)."""
    v = np.ascontiguousarray(verts, np.float32).reshape(-1)
    t = np.ascontiguousarray(tris, np.uint32).reshape(-1)
    nt = t.size // 3
    lo = np.zeros((2 * nt, 4), np.float32)
    hi = np.zeros((2 * nt, 4), np.float32)
    prim = np.zeros(nt, np.uint32)
    root = np.zeros(1, np.uint32)
    lib().orc_lbvh_build(_ptr(v, _fp), _ptr(t, _up), nt, _ptr(lo, _fp), _ptr(hi, _fp), _ptr(prim, _up), _ptr(root, _up))
    return lo, hi, prim, int(root[0])


def ray_tri(p, d, a, b, c):
    arrs = [np.ascontiguousarray(x, np.float32) for x in (p, d, a, b, c)]
    return bool(lib().orc_ray_tri_woop_f(*[_ptr(x, _fp) for x in arrs]))


def ray_hit_lists(verts, tris, pts):
    """Per point, the triangle ids the fixed-direction parity ray hits (all-triangle loop)."""
    v = np.ascontiguousarray(verts, np.float32).reshape(-1)
    t = np.ascontiguousarray(tris, np.uint32).reshape(-1)
    p = np.ascontiguousarray(pts, np.float32).reshape(-1)
    nt, npt = t.size // 3, p.size // 3
    hits = np.zeros(npt, np.int32)
    cap = 8 * npt + 16
    ids = np.zeros(cap, np.uint32)
    tot = lib().orc_ray_hit_lists(_ptr(v, _fp), _ptr(t, _up), nt, _ptr(p, _fp), npt, _ptr(hits, _ip), _ptr(ids, _up), cap)
    assert tot <= cap
    out, o = [], 0
    for h in hits:
        out.append(ids[o:o + h].copy())
        o += h
    return out


def closest_point_triangle(p, a, b, c):
    arrs = [np.ascontiguousarray(x, np.float32) for x in (p, a, b, c)]
    cp = np.zeros(3, np.float32)
    bc = np.zeros(3, np.float32)
    lib().orc_closest_point_triangle_f(*[_ptr(x, _fp) for x in arrs], _ptr(cp, _fp), _ptr(bc, _fp))
    return cp, bc


_ref = None


def ref_lib():
    """The reference's own LBVH code compiled from its sources, or None."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        L = C.CDLL(REF_LIB)
        L.ref_lbvh_build.argtypes = [_fp, C.c_uint32, _up, C.c_uint32, _fp, _fp, _up, _up]
        L.ref_lbvh_refit.argtypes = [_fp, _fp, C.c_uint32, _up, C.c_uint32, _fp, _fp]
        L.ref_query_closest.restype = C.c_int
        L.ref_query_closest.argtypes = [_fp, C.c_uint32, _up, C.c_uint32, _fp, C.c_float, _up, _fp]
        _ref = L
    return _ref


def ref_lbvh_build(verts, tris):
    L = ref_lib()
    v = np.ascontiguousarray(verts, np.float32).reshape(-1)
    t = np.ascontiguousarray(tris, np.uint32).reshape(-1)
    nt = t.size // 3
    lo = np.zeros((2 * nt, 4), np.float32)
    hi = np.zeros((2 * nt, 4), np.float32)
    prim = np.zeros(nt, np.uint32)
    root = np.zeros(1, np.uint32)
    L.ref_lbvh_build(_ptr(v, _fp), v.size // 3, _ptr(t, _up), nt, _ptr(lo, _fp), _ptr(hi, _fp), _ptr(prim, _up),
                     _ptr(root, _up))
    return lo, hi, prim, int(root[0])


def ref_lbvh_refit(verts, new_verts, tris):
    L = ref_lib()
    v = np.ascontiguousarray(verts, np.float32).reshape(-1)
    w = np.ascontiguousarray(new_verts, np.float32).reshape(-1)
    t = np.ascontiguousarray(tris, np.uint32).reshape(-1)
    nt = t.size // 3
    lo = np.zeros((2 * nt, 4), np.float32)
    hi = np.zeros((2 * nt, 4), np.float32)
    L.ref_lbvh_refit(_ptr(v, _fp), _ptr(w, _fp), v.size // 3, _ptr(t, _up), nt, _ptr(lo, _fp), _ptr(hi, _fp))
    return lo, hi


def ref_query_closest(verts, tris, p, max_dist):
    L = ref_lib()
    v = np.ascontiguousarray(verts, np.float32).reshape(-1)
    t = np.ascontiguousarray(tris, np.uint32).reshape(-1)
    pp = np.ascontiguousarray(p, np.float32)
    face = np.zeros(1, np.uint32)
    pt = np.zeros(3, np.float32)
    ok = L.ref_query_closest(_ptr(v, _fp), v.size // 3, _ptr(t, _up), t.size // 3, _ptr(pp, _fp), max_dist,
                             _ptr(face, _up), _ptr(pt, _fp))
    return bool(ok), int(face[0]), pt
