"""``torch.autograd.Function`` over the distributed plan: forward runs
``dfft_execute``, backward runs ``dfft_execute_adjoint`` (both HIP, through
the C ABI).  The reference has no differentiation; this is the harness the
adjoint is checked with.

Gradients follow torch's convention for complex tensors (the vector-Jacobian
product under the real inner product Re sum conj(a) b), which is the adjoint
the library computes.
"""

from __future__ import annotations

import dfft


def _function():
    import torch

    class PlanExecute(torch.autograd.Function):
        @staticmethod
        def forward(ctx, plan, direction, *ins):
            ctx.plan, ctx.direction = plan, int(direction)
            fwd = ctx.direction == int(dfft.Direction.forward)
            shapes = plan.out_extents if fwd else plan.in_extents
            if fwd:
                dt = torch.complex128 if plan.fp64 else torch.complex64
            elif plan.real:
                dt = torch.float64 if plan.fp64 else torch.float32
            else:
                dt = torch.complex128 if plan.fp64 else torch.complex64
            ctx.in_dtypes = [t.dtype for t in ins]
            outs = [torch.empty(s, dtype=dt, device=ins[0].device) for s in shapes]
            plan._execute([t.detach().contiguous() for t in ins], outs, ctx.direction)
            return tuple(outs)

        @staticmethod
        def backward(ctx, *gouts):
            plan = ctx.plan
            fwd = ctx.direction == int(dfft.Direction.forward)
            shapes = plan.in_extents if fwd else plan.out_extents
            grads = [torch.empty(s, dtype=d, device=gouts[0].device) for s, d in zip(shapes, ctx.in_dtypes)]
            g = [t.contiguous() for t in gouts]
            plan.execute_adjoint(g, grads, ctx.direction)
            return (None, None) + tuple(grads)

    return PlanExecute


_PLAN_EXECUTE = None


def execute(plan: "dfft.Plan", ins, direction):
    """Differentiable ``execute``: returns the P output blocks as new tensors."""
    global _PLAN_EXECUTE
    if _PLAN_EXECUTE is None:
        _PLAN_EXECUTE = _function()
    return list(_PLAN_EXECUTE.apply(plan, int(direction), *ins))
