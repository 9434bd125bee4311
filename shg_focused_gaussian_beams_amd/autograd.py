"""torch.autograd wrapper of the SH efficiency, for checking the hand-written adjoint.

``efficiency(ctx, pump)`` returns eta = P_SH(out)/P_FH(z=0) of the plain z-march of
``ctx`` with ``pump`` (complex64 tensor, (ny, nx)) as the z = 0 FH field; backward
calls the library's discrete adjoint (shg_adjoint).  Gradient convention: torch's
for complex inputs, i.e. d eta = Re sum conj(grad) d pump.
"""
import numpy as np
import torch


class _Efficiency(torch.autograd.Function):
    @staticmethod
    def forward(fctx, pump, ctx, signal, checkpoint_interval):
        a = pump.detach().cpu().numpy().astype(np.complex64)
        ctx.set_fields(a, signal)
        res, grad = ctx.adjoint(checkpoint_interval=checkpoint_interval, want_field=True)
        fctx.save_for_backward(torch.from_numpy(grad))
        fctx.device = pump.device
        return torch.tensor(res["eta"], dtype=torch.float64, device=pump.device)

    @staticmethod
    def backward(fctx, gout):
        (grad,) = fctx.saved_tensors
        return (gout.to(torch.complex64).cpu() * grad).to(fctx.device), None, None, None


def efficiency(ctx, pump, signal=None, checkpoint_interval=0):
    if signal is None:
        signal = np.zeros((ctx.params.ny, ctx.params.nx), np.complex64)
    return _Efficiency.apply(pump, ctx, signal, checkpoint_interval)
