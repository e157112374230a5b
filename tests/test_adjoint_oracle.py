"""The FP64 host adjoint of a taped rollout (the checker for the GPU adjoint)
against closed forms and FP64 central finite differences.

Contact dynamics make the loss piecewise smooth with very sharp features: in a
dense, colliding pile the directional slope changes on parameter scales far
below 1e-4, so central differences at eps = 1e-4 average over many of them.
The checks therefore use (a) a gentle scene where eps = 1e-4 resolves the
derivative and (b) the dense scene at an eps where the difference quotient
has converged (two step sizes agree)."""
import numpy as np
import pytest

import oracle_lib as O
import vkxpbd
from vkxpbd import scenes

SDT = 1.0 / 480.0


def make_scene(n1, spacing_r, vel_sigma, seed, r=0.01):
    rng = np.random.default_rng(seed)
    q = scenes.lattice(n1, n1, n1, spacing_r * r, (-0.1, -0.1, 0.02), 0.1 * r, rng)
    v0 = rng.normal(0, vel_sigma, size=q.shape)
    tgt = q.astype(np.float64) + rng.uniform(-0.05, 0.05, size=q.shape)
    return q, v0, tgt, r


class Rollout:
    def __init__(self, q, r, substeps, fp64=True, **cfg):
        self.q, self.r, self.S = q, r, substeps
        self.N = q.shape[0]
        self.o = O.OracleSolver(fp64=fp64)
        self.o.init(scenes.solver_config(vkxpbd, r, **cfg), self.N)
        self.o.finalize_meshes()

    def run(self, v0, relax, tape=False, x0=None):
        q = self.q if x0 is None else x0
        self.o.set_particles(q, np.zeros_like(self.q), np.full(self.N, scenes.PARTICLE_INV_MASS, np.float32),
                             np.full(self.N, self.r, np.float32), np.ones(self.N, np.int32))
        self.o.rollout(v0, relax, SDT, self.S, keep_tape=tape)
        return self.o.get_positions()

    def loss(self, v0, relax, tgt):
        x = self.run(v0, relax)
        return float(np.mean(np.sum((x - tgt) ** 2, axis=1)))

    def grad(self, v0, relax, tgt):
        x = self.run(v0, relax, tape=True)
        return self.o.rollout_backward(2.0 * (x - tgt) / self.N)


def fd(f, x0, k, eps):
    xp = x0.copy().reshape(-1); xp[k] += eps
    xm = x0.copy().reshape(-1); xm[k] -= eps
    return (f(xp.reshape(x0.shape)) - f(xm.reshape(x0.shape))) / (2 * eps)


def test_free_flight_closed_form():
    """No contacts: x_S = x0 + S h v0 + const, so dL/dv0 = S h dL/dx_S and
    dL/dx0 = dL/dx_S; the loss is independent of the relaxation."""
    rng = np.random.default_rng(0)
    q = rng.uniform(-1, 1, (50, 3)).astype(np.float32) * 5 + np.float32([0, 0, 10])
    R = Rollout(q, 0.01, 16, has_ground_plane=0)
    v0 = rng.normal(0, 0.3, q.shape)
    gx = rng.normal(size=q.shape)
    R.run(v0, 0.9, tape=True)
    gx0, gv0, gr = R.o.rollout_backward(gx)
    np.testing.assert_allclose(gx0, gx, rtol=1e-12)
    h = float(np.float32(SDT))   # substep length as the solver holds it
    np.testing.assert_allclose(gv0, 16 * h * gx, rtol=1e-12)
    assert gr == 0.0


def test_head_on_pair_closed_form():
    """Two equal particles approaching along x, no gravity or friction: each
    substep projects the overlap e to e*(1-relax) ... the FD of the exact
    piecewise-linear map equals the adjoint."""
    q = np.float32([[0.0, 0.0, 1.0], [0.0195, 0.0, 1.0]])
    R = Rollout(q, 0.01, 8, has_ground_plane=0, gravity=[0.0, 0.0, 0.0], particle_mu=0.0)
    v0 = np.array([[0.2, 0.0, 0.0], [-0.2, 0.0, 0.0]])
    tgt = np.zeros((2, 3))
    gx0, gv0, gr = R.grad(v0, 0.9, tgt)
    for k in range(6):
        ref = fd(lambda v: R.loss(v, 0.9, tgt), v0, k, 1e-7)
        assert abs(gv0.reshape(-1)[k] - ref) <= 1e-6 * max(1.0, abs(ref)), (k, gv0.reshape(-1)[k], ref)
    ref = (R.loss(v0, 0.9 + 1e-7, tgt) - R.loss(v0, 0.9 - 1e-7, tgt)) / 2e-7
    assert abs(gr - ref) <= 1e-6 * max(1.0, abs(ref))
    assert gr != 0.0


def test_gentle_scene_fd_eps_1e4():
    """Sparse, slow, frictionless pile in free flight (spacing 2.3r, 0.2 m/s):
    every sampled coordinate equals the converged difference quotient to
    1e-3, and at eps = 1e-4 at least 13 of 16 do (a coordinate whose +-1e-4
    interval crosses a contact activation averages two slopes; friction would
    add stick/slip switches)."""
    rng = np.random.default_rng(21)
    q = scenes.lattice(5, 5, 5, 0.023, (-0.1, -0.1, 1.0), 0.001, rng)
    v0 = rng.normal(0, 0.2, size=q.shape)
    tgt = q.astype(np.float64) + rng.uniform(-0.05, 0.05, size=q.shape)
    R = Rollout(q, 0.01, 32, particle_mu=0.0)
    gx0, gv0, gr = R.grad(v0, 0.9, tgt)
    f = lambda v: R.loss(v, 0.9, tgt)
    at_1e4 = 0
    for k in np.random.default_rng(1).choice(v0.size, 16, replace=False):
        ad = gv0.reshape(-1)[k]
        ref = fd(f, v0, k, 1e-7)
        assert abs(ad - ref) <= 1e-3 * max(abs(ref), abs(ad)) + 1e-11, (k, ad, ref)
        r4 = fd(f, v0, k, 1e-4)
        at_1e4 += abs(ad - r4) <= 1e-3 * max(abs(r4), abs(ad)) + 1e-11
    assert at_1e4 >= 13, at_1e4
    ref = (R.loss(v0, 0.9 + 1e-4, tgt) - R.loss(v0, 0.9 - 1e-4, tgt)) / 2e-4
    assert abs(gr - ref) <= 1e-3 * max(abs(ref), abs(gr)), (gr, ref)


def test_dense_scene_fd_converged():
    """Dense colliding pile (lattice spacing 2r, 0.5 m/s): on coordinates where
    the difference quotient has converged (eps 1e-7 and 1e-8 agree to 1e-4),
    it equals the adjoint to 1e-3; the positions' gradient is checked too."""
    q, v0, tgt, r = make_scene(6, 2.0, 0.5, seed=4)
    R = Rollout(q, r, 32)
    gx0, gv0, gr = R.grad(v0, 0.9, tgt)
    rng = np.random.default_rng(2)
    checked = 0
    for k in rng.choice(v0.size, 16, replace=False):
        f = lambda v: R.loss(v, 0.9, tgt)
        a, b = fd(f, v0, k, 1e-7), fd(f, v0, k, 1e-8)
        if abs(a - b) > 1e-4 * max(abs(a), abs(b)) + 1e-12:
            continue
        ad = gv0.reshape(-1)[k]
        assert abs(ad - b) <= 1e-3 * max(abs(b), abs(ad)) + 1e-11, (k, ad, a, b)
        checked += 1
    assert checked >= 10
    x0 = q.astype(np.float64)
    for k in rng.choice(x0.size, 4, replace=False):
        f = lambda x: float(np.mean(np.sum((R.run(v0, 0.9, x0=x.astype(np.float32)) - tgt) ** 2, axis=1)))
        # positions enter as float32: perturb by representable steps
        xp = x0.copy().reshape(-1); xm = x0.copy().reshape(-1)
        xp[k] = np.nextafter(np.float32(xp[k]), np.float32(np.inf))
        xm[k] = np.nextafter(np.float32(xm[k]), np.float32(-np.inf))
        ref = (f(xp.reshape(x0.shape)) - f(xm.reshape(x0.shape))) / (xp[k] - xm[k])
        assert abs(gx0.reshape(-1)[k] - ref) <= 2e-2 * max(abs(ref), 1e-3), (k, gx0.reshape(-1)[k], ref)
