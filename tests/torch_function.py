"""torch.autograd.Function over the forward/backward pair of the C ABI.

Test-only glue: ``InverseDFTLoss.apply(vxc, ctx)`` returns the loss of
``idft_forward`` at the child-mesh potential ``vxc`` (a float64 CPU tensor)
and back-propagates ``idft_backward``'s force, i.e. dL/dv_xc, the gradient the
reference's getForceVector returns (InverseDFTSolverFunction.h:146-148).
Works over any library that implements the ABI (the HIP library or the
oracle).
"""
import numpy as np
import torch


class InverseDFTLoss(torch.autograd.Function):
    @staticmethod
    def forward(fctx, vxc, ctx):
        loss = ctx.forward(vxc.detach().cpu().numpy().astype(np.float64))
        fctx.idft = ctx
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(fctx, grad_out):
        force = torch.from_numpy(fctx.idft.backward())
        return grad_out * force, None
