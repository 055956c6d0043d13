"""kx_forward_noalloc: the forward operator with its scratch in the
library's device array, through the C-ABI, against kx_forward (the same
kernels with the scratch in plan.aux): bitwise equal data on two calls in a
row, the second with a smaller input, and the array reused by a third;
float64 plans and plans whose forward cannot convolve in the spread launch
are rejected with their error codes."""
import ctypes

import numpy as np
import pytest
import torch

from tests import geometry as G

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _op(size, dtype="float32", **kw):
    import pylops_amd
    if size == "large":
        x = np.arange(90) * 10.0
        z = np.arange(70) * 10.0
        t = np.arange(400) * 0.004
        srcs = np.vstack((np.linspace(0, x[-1], 9), np.zeros(9)))
        recs = np.vstack((np.linspace(0, x[-1], 23), np.zeros(23)))
        wav, _, wc = G.ricker(t[:41], f0=20)
        return pylops_amd.Kirchhoff(z, x, t, srcs, recs, 2200.0, wav.astype(dtype), wc,
                                    mode="analytic", dynamic=True, dtype=dtype, **kw)
    # the reference test geometry with nt = 120 ("small") or its nt = 50
    # ("short": the float rows and the filter do not fit the LDS of the
    # sums, so the forward takes two launches and a workspace)
    t = G.t if size == "short" else np.arange(120) * G.PAR["dt"]
    return pylops_amd.Kirchhoff(G.z, G.x, t, G.s2d, G.r2d, G.v0, G.wav.astype(dtype), G.wavc,
                                mode="analytic", dtype=dtype, **kw)


def _noalloc(Op, m):
    from pylops_amd import _lib
    L = _lib.load()
    mt = torch.as_tensor(m, device=DEV)
    d = torch.full((Op.shape[0],), float("nan"), dtype=mt.dtype, device=DEV)
    rc = L.kx_forward_noalloc(ctypes.byref(Op.plan), _lib.ptr(mt), _lib.ptr(d),
                              _lib.stream_ptr(DEV))
    return rc, d


def test_forward_noalloc_matches_forward():
    from pylops_amd import _lib
    rng = np.random.default_rng(3)
    for size in ("large", "small", "large"):
        Op = _op(size)
        m = rng.standard_normal(Op.shape[1]).astype(np.float32)
        rc, d = _noalloc(Op, m)
        _lib.check(rc, "kx_forward_noalloc")
        np.testing.assert_array_equal(d.cpu().numpy(), np.asarray(Op @ m), err_msg=size)


def test_forward_noalloc_rejects_what_needs_a_workspace():
    from pylops_amd import _lib
    L = _lib.load()
    Op = _op("small", dtype="float64")
    rc, _ = _noalloc(Op, np.zeros(Op.shape[1]))
    assert rc == _lib.KX_ETYPE
    assert b"float32 plans only" in L.kx_last_error()
    Op = _op("short")
    rc, _ = _noalloc(Op, np.zeros(Op.shape[1], dtype=np.float32))
    assert rc == _lib.KX_ELDS
    assert b"cannot convolve in the spread launch" in L.kx_last_error()
