"""torch.autograd binding of the GPU eikonal solver.

    values = Eikonal.autograd.eikonal_values(hfmIn, cost)   # cost: leaf tensor on the device
    J = (values * weights).sum(); J.backward()                # cost.grad = dJ/dcost

The forward pass is one solve plus the flow pass that exports the upwind
weights (Solve + PostProcess of HFM_GPU.Interface); the backward pass is one
reverse solve of the linearised scheme with the incoming gradient as
sensitivity (the 'sensitivity' key, _PostProcess.py:194-219), converted from
the reference's costSensitivity convention to a gradient:
dJ/dcost = cost * costSensitivity.  Seed cells carry no cost: their gradient
is 0 (their values are the seed values).
"""
from ._device import torch


class EikonalSolve(torch.autograd.Function):
    """values = u(cost) for the problem described by hfmIn (whose own 'cost'
    and AD keys are ignored)."""

    @staticmethod
    def forward(ctx, cost, hfmIn):
        from . import HFM_GPU  # loads libagd_hip.so
        hfm = hfmIn.copy()
        for key in ('cost', 'speed', 'costVariation', 'sensitivity'):
            if key in hfm:
                del hfm[key]
        hfm['cost'] = cost.detach()
        if str(hfm['model']).startswith('Rander'):
            # the reference's drift-model sensitivities divide by the flow's
            # metric norm (_PostProcess.py:213-214), which is not the
            # gradient of the solve: no gradient is offered for them here
            raise NotImplementedError("EikonalSolve does not differentiate the Rander model")
        iface = HFM_GPU.Interface(hfm)
        iface.keep_flow = True
        iface.SetKernelTraits()
        iface.SetGeometry()
        iface.SetArgs()
        iface.Solve()
        iface.PostProcess()
        if iface.forwardAD:
            raise ValueError("EikonalSolve differentiates through torch; do not pass AD keys")
        values = iface.values.clone()
        ctx.iface = iface
        ctx.save_for_backward(cost)
        return values

    @staticmethod
    def backward(ctx, grad_values):
        (cost,) = ctx.saved_tensors
        iface = ctx.iface
        lam, _ = iface.reverse(grad_values.detach().to(iface.values_t.dtype).unsqueeze(-1))
        grad_cost = lam[..., 0] * cost.detach().to(lam.dtype)
        return grad_cost.to(cost.dtype), None


def eikonal_values(hfmIn, cost):
    """Differentiable values of the eikonal solve with the given cost field
    (a tensor of the grid's shape on the HIP device)."""
    if tuple(cost.shape) != tuple(hfmIn.shape):
        raise ValueError(f"cost must have the grid shape {tuple(hfmIn.shape)}")
    return EikonalSolve.apply(cost, hfmIn)
