"""Convolve1D with a compact filter, computed by the HIP kernel kx_conv
(replaces pylops/signalprocessing/convolve1d.py:57-115, which calls
scipy.signal.fftconvolve(mode="same"))."""
import numpy as np
import torch

from pylops_amd import _lib
from pylops_amd.linearoperator import LinearOperator
from pylops_amd.utils._internal import _value_or_sized_to_tuple
from pylops_amd.utils.backend import like_input, to_device

__all__ = ["Convolve1D", "padded_filter", "conv_rows"]


def padded_filter(h, offset):
    """Zero-pad ``h`` so that index ``offset`` becomes the centre of an
    odd-length filter (convolve1d.py:72-86)."""
    h = np.asarray(h)
    nh = h.size
    if offset > nh - 1:
        raise ValueError("offset must be smaller than h.shape[axis] - 1")
    shift = 2 * (nh // 2 - int(offset))
    if nh % 2 == 0:
        shift -= 1
    if shift > 0:
        h = np.concatenate([np.zeros(shift, dtype=h.dtype), h])
    elif shift < 0:
        h = np.concatenate([h, np.zeros(-shift, dtype=h.dtype)])
    return h


def conv_rows(x, h, out=None):
    """y[r, n] = sum_k h[k] x[r, n + (nh-1)/2 - k] for a contiguous 2-D
    device tensor ``x`` and an odd device filter ``h`` of the same dtype."""
    if out is None:
        out = torch.empty_like(x)
    dt = _lib.KX_F64 if x.dtype == torch.float64 else _lib.KX_F32
    rc = _lib.load().kx_conv(_lib.ptr(x), _lib.ptr(out), x.shape[0], x.shape[1],
                             _lib.ptr(h), h.numel(), dt, _lib.stream_ptr(x.device))
    _lib.check(rc, "kx_conv")
    return out


class Convolve1D(LinearOperator):
    """1D convolution along ``axis`` with a filter shorter than the signal.

    Forward is "same"-mode convolution, adjoint is correlation (convolution
    with the flipped filter), both on the HIP device."""

    def __init__(self, dims, h, offset=0, axis=-1, method=None, dtype="float64",
                 name="C"):
        dims = _value_or_sized_to_tuple(dims)
        h = np.asarray(h.detach().cpu().numpy() if isinstance(h, torch.Tensor) else h)
        if h.ndim != 1:
            raise NotImplementedError("pylops_amd Convolve1D takes a 1-D filter")
        nh = h.size
        if method not in (None, "direct", "fft", "overlapadd"):
            raise NotImplementedError("method must be direct, fft or overlapadd")
        if nh > dims[axis]:
            raise NotImplementedError(
                "filters longer than the signal (_Convolve1Dlong) are not provided")
        super().__init__(dtype=np.dtype(dtype), dims=dims, dimsd=dims, name=name)
        self.axis = axis if axis >= 0 else len(dims) + axis
        self.nh = nh
        self.h = padded_filter(h, offset)
        self.hstar = np.flip(self.h).copy()
        self._dev = {}

    def _filters(self, device, tdtype):
        key = (str(device), tdtype)
        if key not in self._dev:
            self._dev[key] = (torch.as_tensor(self.h, dtype=tdtype, device=device),
                              torch.as_tensor(self.hstar, dtype=tdtype, device=device))
        return self._dev[key]

    def _apply(self, x, adjoint):
        _lib.require_device()
        dev = x.device if isinstance(x, torch.Tensor) and x.is_cuda else torch.device("cuda")
        xd = to_device(x, self.dtype, dev).reshape(self.dims)
        xd = torch.movedim(xd, self.axis, -1).contiguous()
        shp = xd.shape
        h, hstar = self._filters(dev, xd.dtype)
        y = conv_rows(xd.reshape(-1, shp[-1]), hstar if adjoint else h)
        y = torch.movedim(y.reshape(shp), -1, self.axis).reshape(-1)
        return like_input(y, x)

    def _matvec(self, x):
        return self._apply(x, False)

    def _rmatvec(self, x):
        return self._apply(x, True)
