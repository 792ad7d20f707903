"""torch.autograd.Function around the library's adjoint (verification harness).

    J = QGWaveEnergy.apply(psi0, adjoint, BR0, BI0, nsteps)
    J.backward()            # psi0.grad = dJ/dpsi0 from the hand-written adjoint

psi0 is a CPU float64 tensor (n3, n1, n1) of the physical initial
streamfunction; the forward value and the gradient come from one
qg_adj_gradient call (forward run with checkpoints, adjoint sweep on the
GPU).  This wrapper exists so that the adjoint can be checked with torch's
own tools (directional derivatives, composition with other torch code); the
product interface is the C ABI.
"""
from __future__ import annotations

import numpy as np
import torch


class QGWaveEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, psi0: torch.Tensor, adjoint, BR0, BI0, nsteps: int):
        if psi0.dtype != torch.float64 or psi0.device.type != "cpu":
            raise TypeError("psi0 must be a CPU float64 tensor")
        J, grad = adjoint.gradient(psi0.detach().numpy(), np.asarray(BR0), np.asarray(BI0), int(nsteps))
        ctx.save_for_backward(torch.from_numpy(grad))
        return psi0.new_tensor(J)

    @staticmethod
    def backward(ctx, gJ):
        (grad,) = ctx.saved_tensors
        return gJ * grad, None, None, None, None


def wave_energy(psi0: torch.Tensor, adjoint, BR0, BI0, nsteps: int) -> torch.Tensor:
    return QGWaveEnergy.apply(psi0, adjoint, BR0, BI0, nsteps)
