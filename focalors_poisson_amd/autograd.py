"""torch.autograd face of the composite solve: forward = fp_cat_solve,
backward = fp_cat_vjp, both through the C ABI (no torch arithmetic on the
solve path)."""
import torch


class _CompositeSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, solver, ghost_keys, *tensors):
        nd = len(solver.domains)
        fields = [t.detach().clone().contiguous() for t in tensors[:nd]]
        ghost = {k: g.detach().contiguous() for k, g in zip(ghost_keys, tensors[nd:])}
        solver.solve(fields, ghost)
        ctx.solver = solver
        ctx.ghost_keys = ghost_keys
        ctx.like = [(f.shape, f.dtype, f.device) for f in fields]
        return tuple(fields)

    @staticmethod
    def backward(ctx, *grad_u):
        # an output the loss does not use has no gradient: zeros of its shape
        gu = [g if g is not None else torch.zeros(shape, dtype=dt, device=dev)
              for g, (shape, dt, dev) in zip(grad_u, ctx.like)]
        gf, gg = ctx.solver.vjp([g.contiguous() for g in gu], faces=ctx.ghost_keys)
        return (None, None, *gf, *[gg[k].view_as(gg[k]) for k in ctx.ghost_keys])


def composite_solve(solver, fields, ghost=None):
    """Differentiable ConcatPoissonSolver2D.solve: returns the solutions (new
    tensors); gradients flow to `fields` and to the ghost-data tensors."""
    ghost = ghost or {}
    keys = tuple(sorted(ghost))
    return list(_CompositeSolve.apply(solver, keys, *fields, *[ghost[k] for k in keys]))
