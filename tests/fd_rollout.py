"""FP64 rollouts of the 32,768-particle differentiable-rollout pile (64
substeps of 1/480 s, loss = mean squared distance of the final positions to
the seeded targets) for finite differences, one oracle per worker process.
Test helper for tests/test_adjoint_cfg4.py."""
import numpy as np

import oracle_lib as O
import vkxpbd
from vkxpbd import scenes

SDT = 1.0 / 480.0
SUBSTEPS = 64
_state = {}


def scene():
    if "sc" not in _state:
        _state["sc"] = scenes.rollout_pile_32768()
    return _state["sc"]


def _solver():
    if "o" not in _state:
        sc = scene()
        o = O.OracleSolver(fp64=True)
        o.init(scenes.solver_config(vkxpbd, sc["r"]), sc["particles"]["q"].shape[0])
        o.finalize_meshes()
        _state["o"] = o
    return _state["o"]


def run(v0, relax, tape=False):
    """Final positions of the rollout from the scene's positions with
    initial velocities v0 ((N, 3) float64) and contact relaxation relax."""
    p = scene()["particles"]
    o = _solver()
    o.set_particles(p["q"], np.zeros_like(p["q"]), p["inv_mass"], p["radii"], p["flags"])
    o.rollout(v0, relax, SDT, SUBSTEPS, keep_tape=tape)
    return o.get_positions()


def loss_of(v0, relax):
    x = run(v0, relax)
    return float(np.mean(np.sum((x - scene()["target"].astype(np.float64)) ** 2, axis=1)))


def adjoint(relax=0.9):
    """(dL/dv0, dL/drelax) from the FP64 host adjoint of the taped rollout."""
    sc = scene()
    v0 = sc["particles"]["qd"].astype(np.float64)
    x = run(v0, relax, tape=True)
    n = x.shape[0]
    _, gv0, gr = _solver().rollout_backward(2.0 * (x - sc["target"].astype(np.float64)) / n)
    return gv0, gr


def central_difference(job):
    """job = (k, eps, relax): central difference of the loss in velocity
    coordinate k (flattened), or in the relaxation when k < 0."""
    k, eps, relax = job
    v0 = scene()["particles"]["qd"].astype(np.float64)
    if k < 0:
        return (loss_of(v0, relax + eps) - loss_of(v0, relax - eps)) / (2.0 * eps)
    vp = v0.copy().reshape(-1)
    vm = v0.copy().reshape(-1)
    vp[k] += eps
    vm[k] -= eps
    return (loss_of(vp.reshape(v0.shape), relax) - loss_of(vm.reshape(v0.shape), relax)) / (2.0 * eps)
