"""Differentiable rollout of the XPBD particle step as a torch.autograd.Function.

    x, v = rollout(solver, x0, v0, relax, sdt, substeps, checkpoint_every=8)

runs `substeps` substeps of length `sdt` (one iteration each) on the HIP
solver from positions x0 and velocities v0 ((N, 3) float32 CUDA tensors) with
contact relaxation `relax` (a 0-dim tensor or float), and returns the final
positions and velocities.  Backpropagation calls the solver's reverse pass:
hand-written adjoint kernels of integrate, the shape and particle-particle
contact projections (contact set frozen from the forward pass) and the delta
apply, recomputing each checkpoint segment of `checkpoint_every` substeps.
Gradients flow to x0, v0 and relax.  One solver holds one rollout at a time:
a second forward before the first one's backward replaces it.
"""
import torch


class XpbdRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, relax, solver, sdt, substeps, checkpoint_every):
        for name, t in (("x0", x0), ("v0", v0)):
            if not t.is_cuda or t.dtype != torch.float32 or t.dim() != 2 or t.shape[1] != 3:
                raise ValueError(f"rollout: {name} must be a float32 CUDA tensor of shape (N, 3)")
            if t.shape[0] != solver.particle_count:
                raise ValueError(f"rollout: {name} has {t.shape[0]} rows, the solver {solver.particle_count} particles")
        xc = x0.detach().contiguous()
        vc = v0.detach().contiguous()
        xs = torch.empty_like(xc)
        vs = torch.empty_like(xc)
        torch.cuda.synchronize(xc.device)
        solver.rollout_forward(xc.data_ptr(), vc.data_ptr(), float(relax), float(sdt), int(substeps),
                               int(checkpoint_every), xs.data_ptr(), vs.data_ptr())
        ctx.solver = solver
        ctx.relax_like = relax if torch.is_tensor(relax) else None
        return xs, vs

    @staticmethod
    def backward(ctx, gxs, gvs):
        if gxs is None and gvs is None:
            return None, None, None, None, None, None, None
        ref = gxs if gxs is not None else gvs
        gx = gxs.contiguous() if gxs is not None else torch.zeros_like(ref)
        gv = gvs.contiguous() if gvs is not None else None
        gx0 = torch.empty_like(gx)
        gv0 = torch.empty_like(gx)
        torch.cuda.synchronize(gx.device)
        gr = ctx.solver.rollout_backward(gx.data_ptr(), gv.data_ptr() if gv is not None else 0, gx0.data_ptr(),
                                         gv0.data_ptr())
        grelax = None
        if ctx.relax_like is not None:
            grelax = torch.full_like(ctx.relax_like, gr)
        return gx0, gv0, grelax, None, None, None, None


def rollout(solver, x0, v0, relax, sdt, substeps, checkpoint_every=8):
    """Final (positions, velocities) of a differentiable rollout; see module doc."""
    return XpbdRollout.apply(x0, v0, relax, solver, sdt, substeps, checkpoint_every)
