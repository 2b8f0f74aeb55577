"""torch.autograd.Function over the complex128 adjoint sweep of the C ABI.

The reference has no derivative of its sweep (src/waveguide.f90:427-445).  Here the loss of each
realization, L = Tr(t^H t) = sum of its transmission eigenvalues, is a differentiable function of
the disorder potential u(x, y) (units U dx / k, layout (nreal, ny, nxdiso) contiguous, i.e. the
C ABI's (nxdiso, ny, nreal) Fortran order).  The gradient comes from the hand-derived reverse
sweep in the HIP library (rg_adjoint_potential_device), which recomputes the slice Green
functions from checkpoints kept every ceil(sqrt(nx)) slices; nothing is traced.
"""
from __future__ import annotations

import torch

from .binding import PRECISIONS, Plan


class TransmissionLoss(torch.autograd.Function):
    """loss[r] = L(u[r]).  forward runs the adjoint sweep once (loss and dL/du together) and keeps
    dL/du for backward, which scales it by the incoming gradient."""

    @staticmethod
    def forward(ctx, upot: torch.Tensor, plan: Plan, loss: str = "g", precision: str = "fp64"):
        if not upot.is_cuda or upot.dtype != torch.float64:
            raise TypeError("upot must be a CUDA float64 tensor of shape (nreal, ny, nxdiso)")
        if tuple(upot.shape[1:]) != (plan.ny, plan.nxdiso):
            raise ValueError(f"upot shape {tuple(upot.shape)} does not match (nreal, {plan.ny}, {plan.nxdiso})")
        if precision not in ("fp64", "mixed_fp32"):
            raise ValueError("precision must be 'fp64' or 'mixed_fp32'")
        u = upot.detach().contiguous()
        torch.cuda.current_stream(u.device).synchronize()
        plan.set_option("adjoint_precision", PRECISIONS[precision])
        val, grad = plan.adjoint_device(u, loss)
        plan.synchronize()
        ctx.save_for_backward(grad)
        return val

    @staticmethod
    def backward(ctx, gout: torch.Tensor):
        (grad,) = ctx.saved_tensors
        return grad * gout.view(-1, 1, 1), None, None, None


def transmission_loss(upot: torch.Tensor, plan: Plan, loss: str = "g", precision: str = "fp64") -> torch.Tensor:
    """Per-realization L(u) with gradient support (see TransmissionLoss).  precision "mixed_fp32"
    runs the adjoint sweep on complex64 state and FP32 matrix cores."""
    return TransmissionLoss.apply(upot, plan, loss, precision)
