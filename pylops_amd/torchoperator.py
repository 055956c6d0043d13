"""TorchOperator: any pylops_amd operator as a torch.autograd.Function
(pylops/torchoperator.py:23-102, pylops/_torchoperator.py:13-68).

Forward applies ``Op @ x`` and backward ``Op.H @ g``.  For the Kirchhoff
operator both are the HIP kernels (spread + wavelet for the forward,
correlation + gather for the gradient).  Tensors stay on the device: there
is no round trip through numpy or cupy, so ``device`` only selects where a
host input is staged and ``devicetorch`` where outputs are placed."""
import numpy as np
import torch

__all__ = ["TorchOperator"]


class _TorchOperator(torch.autograd.Function):
    """Forward = Op @ x, backward = Op.H @ g (the adjoint kernel)."""

    @staticmethod
    def forward(ctx, x, forw, adj, device, devicetorch):
        ctx.forw = forw
        ctx.adj = adj
        ctx.devicetorch = devicetorch
        ctx.indev = x.device
        y = forw(x.detach())
        return y.to(devicetorch) if devicetorch is not None else y

    @staticmethod
    def backward(ctx, y):
        x = ctx.adj(y.detach())
        return x.to(ctx.indev), None, None, None, None


class TorchOperator:
    """Wrap a pylops_amd operator for torch autograd.

    Parameters follow the reference: ``Op``, ``batch`` (the first axis is a
    batch axis), ``flatten`` (inputs are flat per sample), ``device`` and
    ``devicetorch`` (device of the outputs; default: the input's device)."""

    def __init__(self, Op, batch=False, flatten=True, device="cpu", devicetorch=None):
        self.device = device
        self.devicetorch = devicetorch
        self.dtype = np.dtype(Op.dtype)
        self.dims, self.dimsd = Op.dims, Op.dimsd
        self.name = getattr(Op, "name", None)
        nd = 2 if flatten else len(self.dims) + 1
        self.transpf = tuple(np.roll(np.arange(nd), -1).tolist())
        self.transpb = tuple(np.roll(np.arange(nd), 1).tolist())
        if not batch:
            self.matvec = lambda x: Op @ x
            self.rmatvec = lambda x: Op.H @ x
        else:
            self.matvec = lambda x: (Op @ x.permute(self.transpf)).permute(self.transpb)
            self.rmatvec = lambda x: (Op.H @ x.permute(self.transpf)).permute(self.transpb)
        self.Top = _TorchOperator.apply

    def __call__(self, x):
        return self.apply(x)

    def apply(self, x):
        """Apply the operator to ``x`` (a torch tensor) with autograd."""
        return self.Top(x, self.matvec, self.rmatvec, self.device, self.devicetorch)
