"""Autograd for the iterative solvers: cola/utils/custom_autodiff.py:4-86 on the engine.

``iterative_autograd(bwd)`` wraps a solver ``fn(A, *args, **kwargs)`` whose operator
tensors may require gradients: the forward runs the engine's kernels on the operator's
tensors, and the backward calls ``bwd((params, output), grads, unflatten, *args,
**kwargs)``, which returns the operator holding the parameter gradients (plus the
gradients of other inputs, ignored here as in the reference, custom_autodiff.py:35-37).
As in the reference, only the operator's tensors are inputs of the autograd node; the
right-hand side and the other arguments are not differentiated.

``vjp_derivs(fun, primals, duals)`` is the vector-Jacobian product of a function of the
operator's tensors (cola/backends/torch_fns.py vjp_derivs): it runs ``fun`` with the
engine's kernels switched off, so torch autograd sees A(theta) @ X as torch arithmetic.
"""
from functools import wraps

import torch

from .ops import engine_disabled


def _needs_grad(params):
    return torch.is_grad_enabled() and any(isinstance(p, torch.Tensor) and p.requires_grad for p in params)


def vjp_derivs(fun, primals, duals):
    """d/dtheta <fun(theta), duals> for the tensors `primals` (None where a tensor does not
    take part, e.g. integer index arrays)."""
    theta = [p.detach().requires_grad_(p.is_floating_point()) for p in primals]
    with torch.enable_grad(), engine_disabled():
        out = fun(*theta)
    live = [t for t in theta if t.requires_grad]
    grads = torch.autograd.grad(out, live, grad_outputs=duals.to(out.dtype), allow_unused=True) if live else []
    it = iter(grads)
    return [next(it) if t.requires_grad else None for t in theta]


def iterative_autograd(iterative_bwd):
    """custom_autodiff.py:4-72: a torch autograd rule for an iterative solver."""

    def wrap(iterative_fn):
        @wraps(iterative_fn)
        def with_grad(A, *args, **kwargs):
            par, unflatten = A.flatten()
            if not _needs_grad(par):
                # nothing to differentiate: the same call without the autograd node
                return iterative_fn(A, *args, **kwargs)

            class Iterative(torch.autograd.Function):
                @staticmethod
                def forward(ctx, *params):
                    output = iterative_fn(unflatten(params), *args, **kwargs)
                    ctx.params = params
                    ctx.output = output
                    return output

                @staticmethod
                def backward(ctx, *grads):
                    dA, *_ = iterative_bwd((ctx.params, ctx.output), grads, unflatten, *args, **kwargs)
                    return tuple(dA.flatten()[0]) if dA is not None else tuple(None for _ in ctx.params)

            return Iterative.apply(*par)

        return with_grad

    return wrap


class GradOperator:
    """The gradient of an operator's tensors, shaped like the operator (what unflatten gives
    for a list of per-tensor gradients, some of them None)."""

    def __init__(self, grads):
        self.grads = list(grads)

    def flatten(self):
        return self.grads, None
