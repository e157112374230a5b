"""Differentiable rollout on the GPU (vkxpbd.autograd over the C-ABI reverse
pass) against the FP64 host adjoint of the oracle evaluated on the same fp32
trajectory (the oracle's fp32 forward is bit-identical to the GPU forward).

Tolerance: the GPU reverse pass runs in fp32; its gradient must match the FP64
adjoint to 1e-3 relative in norm and on 16 seeded coordinates (relative to the
gradient's largest component).  The FP64 adjoint itself is checked against
finite differences in test_adjoint_oracle.py."""
import numpy as np
import pytest
import torch

import oracle_lib as O
import vkxpbd
from gpu_util import log
from vkxpbd import scenes
from vkxpbd.autograd import rollout

pytestmark = pytest.mark.gpu

SDT = 1.0 / 480.0


def setup(q, v0, r, substeps, **cfg):
    N = q.shape[0]
    c = scenes.solver_config(vkxpbd, r, **cfg)
    s = vkxpbd.XpbdSolver()
    s.init(c, N)
    s.finalize_meshes()
    inv_m = np.full(N, scenes.PARTICLE_INV_MASS, np.float32)
    rad = np.full(N, r, np.float32)
    s.set_particles(q, v0.astype(np.float32), inv_m, rad, np.ones(N, np.int32))
    o = O.OracleSolver(fp64=False)
    o.init(c, N)
    o.finalize_meshes()
    o.set_particles(q, v0.astype(np.float32), inv_m, rad, np.ones(N, np.int32))
    return s, o


def gpu_grad(s, q, v0, tgt, relax, substeps, every):
    dev = torch.device("cuda:0")
    x0 = torch.tensor(q, dtype=torch.float32, device=dev, requires_grad=True)
    v = torch.tensor(v0, dtype=torch.float32, device=dev, requires_grad=True)
    rl = torch.tensor(relax, dtype=torch.float32, device=dev, requires_grad=True)
    xs, _ = rollout(s, x0, v, rl, SDT, substeps, every)
    t = torch.tensor(tgt, dtype=torch.float32, device=dev)
    loss = ((xs - t) ** 2).sum(dim=1).mean()
    loss.backward()
    return xs.detach().cpu().numpy(), x0.grad.cpu().numpy(), v.grad.cpu().numpy(), float(rl.grad.cpu()), float(loss)


def host_grad(o, xs_gpu, q, v0, tgt, relax, substeps):
    o.rollout(v0.astype(np.float32).astype(np.float64), float(np.float32(relax)), SDT, substeps, keep_tape=True)
    x = o.get_positions()
    np.testing.assert_array_equal(x, xs_gpu.astype(np.float64))   # same trajectory, bit for bit
    gx = (2.0 * (xs_gpu.astype(np.float32) - tgt.astype(np.float32)) / x.shape[0]).astype(np.float32)
    return o.rollout_backward(gx.astype(np.float64))


def compare(name, g, h, rng, tol=1e-3):
    g, h = np.asarray(g, np.float64).reshape(-1), np.asarray(h, np.float64).reshape(-1)
    scale = np.abs(h).max()
    rel = np.linalg.norm(g - h) / max(np.linalg.norm(h), 1e-30)
    idx = rng.choice(h.size, min(16, h.size), replace=False)
    worst = float(np.max(np.abs(g[idx] - h[idx])) / max(scale, 1e-30))
    log(f"  {name}: |g-h|/|h| = {rel:.2e}, worst of 16 coords {worst:.2e} (scale {scale:.3e})")
    assert rel <= tol
    assert worst <= tol


@pytest.mark.parametrize("every", [1, 8])
def test_rollout_gradients_small(vk, every):
    rng = np.random.default_rng(4)
    r = 0.01
    q = scenes.lattice(6, 6, 6, 2 * r, (-0.1, -0.1, 0.02), 0.1 * r, rng)
    v0 = rng.normal(0, 0.5, size=q.shape).astype(np.float32)
    tgt = (q.astype(np.float64) + rng.uniform(-0.05, 0.05, size=q.shape)).astype(np.float32)
    s, o = setup(q, v0, r, 32)
    log(f"small rollout: 216 particles, 32 substeps, checkpoint every {every}")
    xs, gx0, gv0, gr, loss = gpu_grad(s, q, v0, tgt, 0.9, 32, every)
    hx0, hv0, hr = host_grad(o, xs, q, v0, tgt, 0.9, 32)
    compare("dL/dv0", gv0, hv0, rng)
    compare("dL/dx0", gx0, hx0, rng)
    log(f"  dL/drelax gpu {gr:.6e} host {hr:.6e}")
    assert abs(gr - hr) <= 1e-3 * abs(hr)
    # the solver's state after the reverse pass is the forward's final state
    np.testing.assert_array_equal(s.get_positions(), xs)


def test_checkpoint_interval_does_not_change_gradients(vk):
    rng = np.random.default_rng(6)
    r = 0.01
    q = scenes.lattice(8, 8, 4, 2 * r, (-0.1, -0.1, 0.02), 0.1 * r, rng)
    v0 = rng.normal(0, 0.5, size=q.shape).astype(np.float32)
    tgt = q + np.float32(0.01)
    out = []
    for every in (1, 5, 24):
        s, _ = setup(q, v0, r, 24)
        out.append(gpu_grad(s, q, v0, tgt, 0.9, 24, every))
    for o_ in out[1:]:
        np.testing.assert_array_equal(o_[0], out[0][0])
        np.testing.assert_array_equal(o_[2], out[0][2])
        assert o_[3] == out[0][3]


def test_rollout_32768_particles(vk):
    """32,768 particles, 64 substeps of 1/480 s, checkpoints every 8."""
    sc = scenes.rollout_pile_32768()
    p = sc["particles"]
    q, v0, tgt = p["q"], p["qd"], sc["target"]
    s, o = setup(q, v0, sc["r"], 64)
    log("rollout: 32768 particles, 64 substeps, checkpoint every 8")
    xs, gx0, gv0, gr, loss = gpu_grad(s, q, v0, tgt, 0.9, 64, 8)
    hx0, hv0, hr = host_grad(o, xs, q, v0, tgt, 0.9, 64)
    rng = np.random.default_rng(44)
    log(f"  loss {loss:.6e}")
    compare("dL/dv0", gv0, hv0, rng)
    compare("dL/dx0", gx0, hx0, rng)
    log(f"  dL/drelax gpu {gr:.6e} host {hr:.6e}")
    assert abs(gr - hr) <= 1e-3 * abs(hr)


def test_rollout_after_steps_matches_fresh_solver(vk):
    """The solver keeps particle state in cell order, so after a few steps its
    storage order is a permutation of the particle ids.  A rollout started from
    that state must give the same trajectory and gradients, bit for bit, as one
    started on a fresh solver from the same x0 and v0."""
    rng = np.random.default_rng(12)
    r = 0.01
    q = scenes.lattice(8, 8, 4, 2 * r, (-0.1, -0.1, 0.02), 0.1 * r, rng)
    v0 = rng.normal(0, 0.5, size=q.shape).astype(np.float32)
    tgt = q + np.float32(0.01)
    fresh, _ = setup(q, v0, r, 24)
    want = gpu_grad(fresh, q, v0, tgt, 0.9, 24, 8)
    used, _ = setup(q, v0, r, 24)
    used.step(1.0 / 60.0, 8, 1)
    moved = used.get_positions()
    assert not np.array_equal(moved, q)
    log("rollout after 1 frame of 8 substeps: 256 particles, 24 substeps, checkpoint every 8")
    got = gpu_grad(used, q, v0, tgt, 0.9, 24, 8)
    for a, b in zip(got[:3], want[:3]):
        np.testing.assert_array_equal(a, b)
    assert got[3] == want[3]
    assert got[4] == want[4]


def test_steps_after_a_rollout_with_wrap_planes(vk):
    """With wrap planes the cell table is reset from the rows its last build
    wrote, and the reverse pass rebuilds the tables of earlier substeps.
    Frames stepped after a rollout must match, bit for bit, a fresh solver
    stepped from the rollout's final state."""
    rng = np.random.default_rng(21)
    r = 0.01
    q = scenes.lattice(8, 8, 4, 2 * r, (-0.1, -0.1, 0.02), 0.1 * r, rng)
    v0 = rng.normal(0, 0.5, size=q.shape).astype(np.float32)
    tgt = q + np.float32(0.01)
    a, _ = setup(q, v0, r, 24)
    a.set_wrap_planes(True)
    assert a.wrap_planes
    log("rollout with wrap planes: 256 particles, 24 substeps, then 3 frames")
    xs = gpu_grad(a, q, v0, tgt, 0.9, 24, 8)[0]
    x1, v1 = a.get_positions(), a.get_velocities()
    np.testing.assert_array_equal(x1, xs)
    b, _ = setup(x1, v1, r, 24)
    b.set_wrap_planes(True)
    for f in range(3):
        for s in (a, b):
            s.step(1.0 / 60.0, 8, 1)
        np.testing.assert_array_equal(a.get_positions(), b.get_positions(), err_msg=f"frame {f}")
        np.testing.assert_array_equal(a.get_velocities(), b.get_velocities(), err_msg=f"frame {f}")
