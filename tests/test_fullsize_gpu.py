"""The operator at the benchmark sizes, through size-independent properties
(the oracle restates the reference kernels at small sizes; here the float64
HIP operator, checked against the oracle to 1e-10 in test_kirchhoff_gpu.py,
is the FP64 reference):

* 2D dynamic, 601x300 grid, 61 sources x 301 receivers, nt=1501: float32
  forward and adjoint within 1e-5 relative L2 of float64, adjointness by
  the dot test to 1e-5, linearity;
* 3D byot, 128x128x120 grid, 25 x 400 traces, nt=1001: float32 tables
  within 1e-5 and float16 offset tables within 1e-3 of float64 (forward
  and adjoint), dot test of the float16-table operator to 1e-5.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _rel(a, b):
    a = torch.as_tensor(a, device=DEV).double().reshape(-1)
    b = torch.as_tensor(b, device=DEV).double().reshape(-1)
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def _dot(Op, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    tdt = torch.float64 if Op.dtype == np.float64 else torch.float32
    u = torch.randn(Op.shape[1], generator=g).to(DEV, tdt)
    v = torch.randn(Op.shape[0], generator=g).to(DEV, tdt)
    yy = float(torch.dot(torch.as_tensor(Op.matvec(u)).double(), v.double()))
    xx = float(torch.dot(u.double(), torch.as_tensor(Op.rmatvec(v)).double()))
    return abs(xx - yy) / max(abs(xx), abs(yy))


def test_2d_dynamic_fullsize():
    import bench
    import pylops_amd
    kw, refl, _ = bench.workload_2d(np.random.default_rng(1001))
    Op32 = pylops_amd.Kirchhoff(dtype="float32", **kw)
    kw64 = dict(kw, wav=np.asarray(kw["wav"], dtype=np.float64))
    Op64 = pylops_amd.Kirchhoff(dtype="float64", **kw64)
    m = torch.as_tensor(refl.ravel(), device=DEV)
    d64 = Op64.matvec(m)
    d32 = Op32.matvec(m.float())
    assert _rel(d32, d64) < 1e-5
    assert _rel(Op32.rmatvec(d64.float()), Op64.rmatvec(d64)) < 1e-5
    assert _dot(Op32, 1) < 1e-5
    assert _dot(Op64, 2) < 1e-12
    g = torch.Generator(device="cpu").manual_seed(3)
    m2 = torch.randn(Op32.shape[1], generator=g).to(DEV)
    lhs = Op32.matvec(m.float() + 2.0 * m2)
    rhs = d32 + 2.0 * Op32.matvec(m2)
    assert _rel(lhs, rhs) < 1e-6


def test_3d_byot_fullsize_fp16_tables():
    import bench
    import pylops_amd
    kw, refl, _ = bench.workload_3d(np.random.default_rng(1002), DEV)
    m = torch.as_tensor(refl.ravel(), device=DEV)
    kw64 = dict(kw, wav=np.asarray(kw["wav"], dtype=np.float64))
    Op64 = pylops_amd.Kirchhoff(dtype="float64", **kw64)
    d64 = Op64.matvec(m)
    a64 = Op64.rmatvec(d64)
    del Op64
    torch.cuda.empty_cache()
    for tabs, bar in ((None, 1e-5), ("float16", 1e-3)):
        Op = pylops_amd.Kirchhoff(dtype="float32", tabledtype=tabs, **kw)
        assert _rel(Op.matvec(m.float()), d64) < bar, tabs
        assert _rel(Op.rmatvec(d64.float()), a64) < bar, tabs
        if tabs == "float16":
            assert _dot(Op, 4) < 1e-5
        del Op
        torch.cuda.empty_cache()
