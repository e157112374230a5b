"""The drop-in boundary without a GPU: libvkxpbd.so loads and exports every
symbol include/xpbd_capi.h declares; the Python module has the reference's
surface (src/bindings.cpp:11-300); host-side errors carry the reference's
messages."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xpbd_capi.h")
LIB = os.path.join(ROOT, "newton-xpbd-vulkan_amd", "vkxpbd", "libvkxpbd.so")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(xpbd_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    names = declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_config_defaults_match_reference():
    import vkxpbd
    c = vkxpbd.SolverConfig()
    want = dict(dt=1.0 / 60.0, substeps=1, iterations=2, particle_v_max=1e5, soft_contact_mu=0.5,
                shape_material_mu=1.0, soft_contact_relaxation=0.9, particle_adhesion=0.0, particle_mu=0.5,
                particle_cohesion=0.0, particle_max_radius=0.02, search_radius=0.04, soft_contact_margin=0.01,
                enable_restitution=False, soft_contact_restitution=0.0, has_ground_plane=True,
                ground_plane_altitude=0.0, shader_dir="shaders", device_index=-1, compat_mode=False,
                enable_validation=False, enable_diagnostics=False, enable_profile=False, max_diag_pairs=100000,
                substep_batch_size=15)
    for k, v in want.items():
        got = getattr(c, k)
        if isinstance(v, float):
            assert got == pytest.approx(v, rel=1e-6), k
        else:
            assert got == v and type(got) is type(v), k
    assert c.gravity == pytest.approx([0.0, 0.0, -9.81])
    c.gravity = (1.0, 2.0, 3.0, 4.0)      # any sequence of length >= 3
    assert c.gravity == [1.0, 2.0, 3.0]
    c.gravity = [5.0]                      # shorter: ignored, as bindings.cpp:41-44
    assert c.gravity == [1.0, 2.0, 3.0]


def test_solver_surface():
    import vkxpbd
    s = vkxpbd.XpbdSolver()
    for m in ("init", "add_mesh", "finalize_meshes", "set_particles", "step", "get_positions", "get_velocities",
              "collide_only", "step_solve_only", "update_mesh", "get_contact_particles", "get_contact_shapes",
              "get_contact_normals", "get_contact_body_positions", "get_contact_body_velocities",
              "set_contact_data", "get_sorted_cells", "get_sorted_point_ids", "get_cell_starts", "get_cell_ends",
              "get_deltas", "get_diag_pair_count", "get_diag_overflow_count", "reset_diag_buffers",
              "get_diag_pairs", "get_p11_diag", "get_p11_apply_diag", "get_shape_contact_diag",
              "get_mesh_candidate_diag", "get_mesh_bvh_lowers", "get_mesh_bvh_uppers", "print_profiler_summary"):
        assert callable(getattr(s, m)), m
    assert s.particle_count == 0 and s.mesh_count == 0 and s.last_contact_count == 0
    assert isinstance(vkxpbd.enumerate_devices(), list)


def test_host_errors_before_init():
    import vkxpbd
    s = vkxpbd.XpbdSolver()
    q = np.zeros((4, 3), np.float32)
    with pytest.raises(RuntimeError, match="XpbdSolver not initialized"):
        s.set_particles(q)
    with pytest.raises(RuntimeError, match="XpbdSolver::init must be called before finalize_meshes"):
        s.finalize_meshes()
    with pytest.raises(RuntimeError, match=r"update_mesh: call finalize_meshes\(\) first"):
        s.update_mesh(0, q)
    with pytest.raises(RuntimeError, match="update_mesh: velocities size mismatch with vertices"):
        s.update_mesh(0, q, np.zeros(6, np.float32))


def test_init_without_device_fails_loudly():
    import vkxpbd
    if vkxpbd.enumerate_devices():
        pytest.skip("a GPU is visible")
    s = vkxpbd.XpbdSolver()
    with pytest.raises(RuntimeError, match="No HIP devices found"):
        s.init(vkxpbd.SolverConfig(), 16)


def test_add_mesh_watertight_detection_host_only():
    """add_mesh runs on the host (scene.cpp:154-202); mesh_count follows it."""
    import vkxpbd
    from vkxpbd.scenes import icosphere
    s = vkxpbd.XpbdSolver()
    v, t = icosphere(1)
    s.add_mesh(v.reshape(-1), t.reshape(-1))
    s.add_mesh(v.reshape(-1), t[:-1].reshape(-1))
    assert s.mesh_count == 2


def test_sort_pairs_rejects_bad_arguments_without_a_gpu():
    """xpbd_sort_pairs validates its arguments before any device work and
    reports through xpbd_last_error(NULL)."""
    lib = C.CDLL(LIB)
    lib.xpbd_last_error.restype = C.c_char_p
    lib.xpbd_sort_pairs.argtypes = [C.c_void_p] * 4 + [C.c_size_t, C.c_int, C.c_void_p]
    assert lib.xpbd_sort_pairs(None, None, None, None, 16, 0, None) == 1
    assert b"key_bits" in lib.xpbd_last_error(None)
    assert lib.xpbd_sort_pairs(None, None, None, None, 16, 24, None) == 1
    assert b"null buffer" in lib.xpbd_last_error(None)
    assert lib.xpbd_sort_pairs(None, None, None, None, 0, 24, None) == 0   # nothing to sort


def test_init_batch_rejects_2049_scenes_without_a_gpu():
    """A batched solver takes at most 2,048 scenes: the batched key holds the
    scene in the 11 bits above the 21-bit cell.  The check comes before any
    device work."""
    import vkxpbd
    s = vkxpbd.XpbdSolver()
    with pytest.raises(RuntimeError, match="at most 2048 scenes"):
        s.init_batch(vkxpbd.SolverConfig(), np.ones(2049, np.int64))
