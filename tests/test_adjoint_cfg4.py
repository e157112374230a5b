"""The FP64 host adjoint against FP64 central finite differences on the
differentiable-rollout workload itself: 32,768 particles (32^3 lattice,
r = 0.01, N(0, 0.5) initial velocities), 64 substeps of 1/480 s, loss = mean
squared distance to the seeded targets, gradients in 16 seeded initial
velocity coordinates and in the contact relaxation.

This dense, colliding pile makes the loss piecewise smooth with features far
below eps = 1e-4: at that step the difference quotients average over many of
them and agree with no exact gradient (the count is printed, not asserted).
The check is therefore made on a ladder of steps (1e-7 ... 1e-12): wherever
two consecutive steps give the same quotient to 1e-3, the difference quotient
has converged, and the adjoint must equal it to 1e-3.  Slow (about a minute
on eight cores): the rollouts run in worker processes."""
import multiprocessing as mp
import os

import numpy as np

import fd_rollout as F

LADDER = [1e-7, 1e-8, 1e-9, 1e-10, 1e-11, 1e-12]


def rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def test_adjoint_vs_central_differences_on_rollout_pile_32768():
    gv0, gr = F.adjoint(0.9)
    nv = gv0.size
    coords = [int(k) for k in np.random.default_rng(4).choice(nv, 16, replace=False)] + [-1]
    ad = np.array([gv0.reshape(-1)[k] for k in coords[:-1]] + [gr])
    epss = [1e-4] + LADDER
    jobs = [(k, e, 0.9) for k in coords for e in epss]
    workers = max(1, min(8, os.cpu_count() or 1))
    os.environ["ORC_THREADS"] = "1"   # one oracle thread per worker process
    with mp.get_context("spawn").Pool(workers) as pool:
        fd = np.array(pool.map(F.central_difference, jobs, chunksize=1)).reshape(len(coords), len(epss))
    at_1e4 = sum(rel(ad[j], fd[j, 0]) <= 1e-3 for j in range(len(coords)))
    print(f"\neps 1e-4: {at_1e4}/{len(coords)} coordinates agree with the adjoint to 1e-3")
    checked = 0
    for j, k in enumerate(coords):
        ladder = fd[j, 1:]
        conv = [m for m in range(len(LADDER) - 1) if rel(ladder[m], ladder[m + 1]) <= 1e-3]
        name = "relax" if k < 0 else f"v0[{k}]"
        print(f"{name:>10}: adjoint {ad[j]: .6e}  fd " + " ".join(f"{v: .4e}" for v in fd[j]) +
              (f"  converged at {LADDER[conv[0] + 1]:g}" if conv else "  not converged"))
        if conv:
            m = conv[0] + 1
            assert rel(ad[j], ladder[m]) <= 1e-3, (name, ad[j], ladder[m])
            checked += 1
    print(f"converged coordinates: {checked}/{len(coords)}, all equal to the adjoint to 1e-3")
    assert checked >= 8
