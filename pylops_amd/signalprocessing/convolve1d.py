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


def _check_method(method, dims):
    """The method rule of _choose_convfunc (C1D:21-47): 1-D signals take
    "direct" (default) or "fft", N-D signals "fft" (default) or
    "overlapadd".  Every method runs the same device kernels here."""
    if len(dims) == 1:
        method = "direct" if method is None else method
        if method not in ("direct", "fft"):
            raise NotImplementedError("method must be direct or fft")
    else:
        method = "fft" if method is None else method
        if method not in ("fft", "overlapadd"):
            raise NotImplementedError("method must be fft or overlapadd")
    return method


def _offset_shift(n, offset):
    """2 (n//2 - offset), minus one for even n (C1D:78-80, 139-141)."""
    shift = 2 * (n // 2 - int(offset))
    if n % 2 == 0:
        shift -= 1
    return shift


def _filter_rows(h, axis, dims):
    """An N-D filter as one row per trace: the filter axis last, the other
    axes broadcast to the signal's (each of size 1 or the signal's)."""
    if h.ndim != len(dims):
        raise ValueError(f"a {h.ndim}-D filter for a {len(dims)}-D signal")
    for k, (a, b) in enumerate(zip(h.shape, dims)):
        if k != axis and a not in (1, b):
            raise ValueError(f"filter shape {h.shape} does not broadcast to {dims} "
                             f"off axis {axis}")
    hm = np.moveaxis(h, axis, -1)
    other = tuple(d for k, d in enumerate(dims) if k != axis)
    return np.ascontiguousarray(np.broadcast_to(hm, other + (hm.shape[-1],))
                                .reshape(-1, hm.shape[-1]))


class Convolve1D(LinearOperator):
    """1D convolution along ``axis`` (C1D:208-323), on the HIP device.

    A filter no longer than the signal (C1D:57-115): forward is "same"-mode
    convolution, adjoint is correlation (convolution with the flipped
    filter).  ``h`` is 1-D (one filter) or has the signal's number of
    dimensions (one filter per trace, other axes of size 1 or the
    signal's).  A filter longer than the signal (_Convolve1Dlong,
    C1D:118-199): the data has the filter's shape, the forward convolves the
    padded model with the filter ("same" to the filter) and the adjoint
    takes the centre lags of the full correlation."""

    def __init__(self, dims, h, offset=0, axis=-1, method=None, dtype="float64",
                 name="C"):
        dims = _value_or_sized_to_tuple(dims)
        h = np.asarray(h.detach().cpu().numpy() if isinstance(h, torch.Tensor) else h)
        axis = axis if axis >= 0 else len(dims) + axis
        if not 0 <= axis < len(dims):
            raise ValueError(f"axis {axis} out of range for {len(dims)}-D dims")
        nh = h.size if h.ndim == 1 else h.shape[axis]
        self.method = _check_method(method, dims)
        self.long = nh > dims[axis]
        self.axis = axis
        self.nh = nh
        self._dev = {}
        if not self.long:
            super().__init__(dtype=np.dtype(dtype), dims=dims, dimsd=dims, name=name)
            if h.ndim == 1:
                self.h = padded_filter(h, offset)
                self.hstar = np.flip(self.h).copy()
                self._rows = False
            else:
                if offset > nh - 1:
                    raise ValueError("offset must be smaller than h.shape[axis] - 1")
                rows = _filter_rows(h, axis, dims)
                self.h = np.stack([padded_filter(r, offset) for r in rows])
                # flipped along the filter axis (the reference flips the last
                # axis, which is the filter axis when axis = -1)
                self.hstar = np.ascontiguousarray(np.flip(self.h, axis=-1))
                self._rows = True
            return
        # filter longer than the signal: the data takes the filter's shape
        if h.ndim == 1 and len(dims) > 1:
            raise ValueError("a 1-D filter longer than the signal needs a 1-D signal "
                             "(the data takes the filter's shape)")
        if h.ndim > 1:
            if h.ndim != len(dims) or any(a != b for k, (a, b) in enumerate(zip(h.shape, dims))
                                          if k != axis):
                raise ValueError(f"filter shape {h.shape} must match dims {dims} off axis "
                                 f"{axis}")
        if offset > dims[axis] - 1:
            raise ValueError("offset must be smaller than self.dims[self.axis] - 1")
        super().__init__(dtype=np.dtype(dtype), dims=dims, dimsd=h.shape, name=name)
        self.h = np.ascontiguousarray(np.moveaxis(h, axis, -1).reshape(-1, nh)) \
            if h.ndim > 1 else h.reshape(1, nh)
        self.hstar = None
        self._shift = _offset_shift(dims[axis], offset)
        self._rows = h.ndim > 1

    def _filters(self, device, tdtype):
        key = (str(device), tdtype)
        if key not in self._dev:
            self._dev[key] = (
                torch.as_tensor(self.h, dtype=tdtype, device=device),
                None if self.hstar is None else torch.as_tensor(self.hstar, dtype=tdtype,
                                                                device=device))
        return self._dev[key]

    def _short(self, xd, adjoint):
        h, hstar = self._filters(xd.device, xd.dtype)
        f = hstar if adjoint else h
        xr = xd.reshape(-1, xd.shape[-1])
        if not self._rows:
            return conv_rows(xr, f)
        y = torch.empty_like(xr)
        dt = _lib.KX_F64 if xr.dtype == torch.float64 else _lib.KX_F32
        rc = _lib.load().kx_conv_rows(_lib.ptr(xr), xr.shape[1], _lib.ptr(y), xr.shape[0],
                                      xr.shape[1], _lib.ptr(f), f.shape[1], f.shape[1], dt,
                                      _lib.stream_ptr(xr.device))
        _lib.check(rc, "kx_conv_rows")
        return y

    def _long_forward(self, xd):
        """Rows of the model (n samples) -> rows of the data (nh samples):
        convolve(h, pad(x), "same") (C1D:165-174) as a "same" convolution
        of the long filter row with the padded model row as the filter (a
        zero in front when its length is even, to centre it)."""
        h, _ = self._filters(xd.device, xd.dtype)
        xr = xd.reshape(-1, xd.shape[-1])
        ntr, n = xr.shape
        sh = self._shift
        lead = max(sh, 0)
        m = n + abs(sh)
        if m % 2 == 0:
            lead += 1
        nf = lead + n + max(-sh, 0)
        f = torch.zeros((ntr, nf), dtype=xr.dtype, device=xr.device)
        f[:, lead:lead + n] = xr
        y = torch.empty((ntr, self.nh), dtype=xr.dtype, device=xr.device)
        dt = _lib.KX_F64 if xr.dtype == torch.float64 else _lib.KX_F32
        rc = _lib.load().kx_conv_rows(_lib.ptr(h), self.nh if self._rows else 0, _lib.ptr(y),
                                      ntr, self.nh, _lib.ptr(f), nf, nf, dt,
                                      _lib.stream_ptr(xr.device))
        _lib.check(rc, "kx_conv_rows")
        return y

    def _long_adjoint(self, dd):
        """Rows of the data (nh) -> rows of the model (n): the full
        convolution of the padded data with the flipped filter, centre n
        samples (C1D:176-199), as a correlation with the filter at n lags."""
        h, _ = self._filters(dd.device, dd.dtype)
        dr = dd.reshape(-1, self.nh)
        ntr = dr.shape[0]
        n = self.dims[self.axis]
        sh = self._shift
        padd0 = max(-sh, 0)
        full = self.nh + (self.nh + abs(sh)) - 1
        s0 = full // 2 - n // 2
        lag0 = s0 - (self.nh - 1) - padd0
        y = torch.empty((ntr, n), dtype=dr.dtype, device=dr.device)
        dt = _lib.KX_F64 if dr.dtype == torch.float64 else _lib.KX_F32
        rc = _lib.load().kx_xcorr_rows(_lib.ptr(h), self.nh if self._rows else 0, self.nh,
                                       _lib.ptr(dr), self.nh, self.nh, _lib.ptr(y), ntr, n,
                                       lag0, dt, _lib.stream_ptr(dr.device))
        _lib.check(rc, "kx_xcorr_rows")
        return y

    def _apply(self, x, adjoint):
        _lib.require_device()
        dev = x.device if isinstance(x, torch.Tensor) and x.is_cuda else torch.device("cuda")
        shape_in = self.dimsd if adjoint else self.dims
        xd = to_device(x, self.dtype, dev).reshape(shape_in)
        xd = torch.movedim(xd, self.axis, -1).contiguous()
        shp = xd.shape
        if not self.long:
            y = self._short(xd, adjoint)
        elif adjoint:
            y = self._long_adjoint(xd)
        else:
            y = self._long_forward(xd)
        out_last = self.dims[self.axis] if adjoint else self.dimsd[self.axis]
        y = torch.movedim(y.reshape(shp[:-1] + (out_last,)), -1, self.axis).reshape(-1)
        return like_input(y, x)

    def _matvec(self, x):
        return self._apply(x, False)

    def _rmatvec(self, x):
        return self._apply(x, True)
