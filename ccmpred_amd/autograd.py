"""The HIP PLL objective as a torch.autograd.Function.

``pll_objective(x, problem)`` returns f(x) as a 0-d FP64 tensor; its
backward is the library's hand-written adjoint and yields ``grad_f * g`` with
g the reference's *tied* gradient (evaluate_cpu.c:131-161): for a coupling
W(b,k,a,j) with k != j, g is the derivative of f when W(b,k,a,j) and its
mirror W(a,j,b,k) move together (numdif.c:29-45); self-edge and padding
entries get only their L2 term.  It equals the ordinary gradient for the
single potentials.

Two ways to get the gradient:
  * ``recompute=True`` (default): the forward computes f only and keeps no
    per-sequence probabilities; the backward recomputes the logits and
    probabilities chunk by chunk while it accumulates the pair gradient.
  * ``recompute=False``: the forward keeps the N x 21L probability operand on
    the device (needs an unchunked problem) and the backward only runs the
    pair-gradient contraction and the assembly.
"""
from __future__ import annotations

import torch


class PLLObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, problem, recompute=True):
        if recompute:
            f = problem.objective(x)
        else:
            f = problem.forward(x)
        ctx.problem = problem
        ctx.recompute = recompute
        ctx.save_for_backward(x)
        return torch.tensor(f, dtype=torch.float64, device=x.device)

    @staticmethod
    def backward(ctx, grad_out):
        (x,) = ctx.saved_tensors
        problem = ctx.problem
        if ctx.recompute:
            _, g = problem.evaluate(x)
        else:
            g = problem.backward(x)
        return (grad_out.to(torch.float32) * g), None, None


def pll_objective(x: torch.Tensor, problem, recompute: bool = True) -> torch.Tensor:
    return PLLObjective.apply(x, problem, recompute)
