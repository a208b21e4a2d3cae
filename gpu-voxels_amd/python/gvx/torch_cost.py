"""torch.autograd.Function over the hand-written clearance-cost kernels.

The forward and backward passes are the C-ABI entry points gvx_cost_forward /
gvx_cost_backward; torch only provides device memory and the autograd tape.
"""
import torch

from . import Cost


def _sync():
    torch.cuda.current_stream().synchronize()


class ClearanceCostFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, cost):
        q = q.detach().contiguous().float()
        out = torch.empty(q.shape[0], device=q.device, dtype=torch.float32)
        _sync()
        cost.forward(q, out)
        ctx.cost = cost
        ctx.save_for_backward(q)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        (q,) = ctx.saved_tensors
        up = grad_out.detach().contiguous().float()
        dq = torch.empty_like(q)
        _sync()
        ctx.cost.backward(q, dq, None, up)
        return dq, None


def clearance_cost(q, cost: Cost):
    """q: (T, W, L) float32 CUDA tensor of joint values -> (T,) cost per trajectory."""
    return ClearanceCostFn.apply(q, cost)
