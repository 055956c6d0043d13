"""TorchOperator / autograd on the HIP operator
(pytests/test_torchoperator.py:28-103 restated), plus a central
finite-difference check of the gradient of 0.5 |K m - d|^2."""
import numpy as np
import pytest

from tests import geometry as G
from tests.test_solvers_gpu import _TorchDense

pytestmark = pytest.mark.gpu


def _kop(dtype):
    import pylops_amd
    return pylops_amd.Kirchhoff(G.z, G.x, G.t, G.s2d, G.r2d, G.v0, G.wav, G.wavc,
                                mode="analytic", dynamic=True, dtype=dtype)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_torchoperator_forward_and_gradient(dtype):
    import torch
    from pylops_amd import TorchOperator
    Op = _kop(dtype)
    Top = TorchOperator(Op, batch=False, device="gpu")
    rng = np.random.default_rng(0)
    x = rng.standard_normal(Op.shape[1]).astype(dtype)
    v = rng.standard_normal(Op.shape[0]).astype(dtype)
    xt = torch.as_tensor(x, device="cuda").view(-1)
    xt.requires_grad = True
    vt = torch.as_tensor(v, device="cuda").view(-1)
    yt = Top.apply(xt)
    yt.backward(vt, retain_graph=True)
    tol = 1e-6 if dtype == "float32" else 1e-14
    # forward / gradient are the operator's own kernels (LDS float64
    # accumulation order may change the last float32 bit between calls)
    np.testing.assert_allclose(yt.detach().cpu().numpy(), Op @ x, rtol=tol,
                               atol=tol * np.abs(Op @ x).max())
    np.testing.assert_allclose(xt.grad.cpu().numpy(), Op.H @ v, rtol=tol,
                               atol=tol * np.abs(Op.H @ v).max())


def test_torchoperator_batch_dense():
    import torch
    from pylops_amd import TorchOperator
    rng = np.random.default_rng(1)
    A = rng.normal(0.0, 1.0, (11, 11)).astype(np.float32)
    Dop = _TorchDense.make(A)
    Top = TorchOperator(Dop, batch=True, device="gpu")
    x = rng.normal(0.0, 1.0, (4, 11)).astype(np.float32)
    xt = torch.as_tensor(x, device="cuda")
    xt.requires_grad = True
    yt = Top.apply(xt)
    np.testing.assert_allclose(yt.detach().cpu().numpy(), (A.astype(np.float64) @ x.T).T,
                               rtol=1e-6, atol=1e-6)
    yt.sum().backward()
    np.testing.assert_allclose(xt.grad.cpu().numpy(),
                               np.tile(A.astype(np.float64).sum(0), (4, 1)), rtol=1e-5)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_gradient_finite_differences(dtype):
    """loss = 0.5 |K m - d|^2: the gradient from the adjoint kernel matches
    central differences (eps = 1e-3) along 32 random unit directions.  The
    loss is quadratic, so central differences are exact up to rounding."""
    import torch
    from pylops_amd import TorchOperator
    Op = _kop(dtype)
    Top = TorchOperator(Op)
    tdt = torch.float64 if dtype == "float64" else torch.float32
    rng = np.random.default_rng(2)
    d = torch.as_tensor(rng.standard_normal(Op.shape[0]), dtype=tdt, device="cuda")
    m0 = torch.as_tensor(rng.standard_normal(Op.shape[1]), dtype=tdt, device="cuda")
    m = m0.clone().requires_grad_(True)
    loss = 0.5 * ((Top.apply(m) - d) ** 2).sum()
    loss.backward()
    g = m.grad.double()
    eps = 1e-3
    errs = []
    for k in range(32):
        u = torch.as_tensor(rng.standard_normal(Op.shape[1]), dtype=torch.float64, device="cuda")
        u /= u.norm()
        with torch.no_grad():
            lp = 0.5 * ((Top.apply((m0.double() + eps * u).to(tdt)).double() - d.double()) ** 2).sum()
            lm = 0.5 * ((Top.apply((m0.double() - eps * u).to(tdt)).double() - d.double()) ** 2).sum()
        fd = float((lp - lm) / (2 * eps))
        an = float((g * u).sum())
        errs.append(abs(fd - an) / max(abs(an), 1e-12 * float(g.norm())))
    tol = 1e-6 if dtype == "float64" else 2e-2
    assert np.median(errs) < tol, errs


def test_matmat_columns_are_the_matvecs():
    """Kirchhoff.matmat / rmatmat (LO:436-447: column by column): every
    column is bitwise its matvec / rmatvec, and the counters move as the
    reference's do (one matmat, no net matvec)."""
    Op = _kop("float32")
    rng = np.random.default_rng(7)
    X = rng.standard_normal((Op.shape[1], 5)).astype(np.float32)
    V = rng.standard_normal((Op.shape[0], 5)).astype(np.float32)
    Op.reset_count()
    Y = Op.matmat(X)
    Z = Op.rmatmat(V)
    assert (Op.matmat_count, Op.rmatmat_count, Op.matvec_count, Op.rmatvec_count) == (1, 1, 0, 0)
    assert Y.shape == (Op.shape[0], 5) and Z.shape == (Op.shape[1], 5)
    for j in range(5):
        np.testing.assert_array_equal(Y[:, j], Op.matvec(X[:, j]))
        np.testing.assert_array_equal(Z[:, j], Op.rmatvec(V[:, j]))
    # table reassignment reaches matmat
    Op.trav_recs = Op.trav_recs * 1.01
    Y2 = Op.matmat(X)
    np.testing.assert_array_equal(Y2[:, 1], Op.matvec(X[:, 1]))
    assert G.relerr(Y2, Y) > 1e-3


def test_torchoperator_batch_kirchhoff_vs_oracle():
    """pytests/test_torchoperator.py:61-80 on the Kirchhoff operator: the
    batched forward equals matmat exactly; the gradient equals rmatmat
    exactly and the oracle's adjoint (the reference's TorchOperator
    gradient is its adjoint) to 1e-5."""
    import torch
    from oracle import oracle as O
    from pylops_amd import TorchOperator
    Op = _kop("float32")
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav,
              wavcenter=G.wavc, mode="analytic", dynamic=True, dtype="float32")
    orc = O.KirchhoffOracle(**kw)
    Top = TorchOperator(Op, batch=True, device="gpu")
    rng = np.random.default_rng(8)
    x = rng.standard_normal((4, Op.shape[1])).astype(np.float32)
    v = rng.standard_normal((4, Op.shape[0])).astype(np.float32)
    xt = torch.as_tensor(x, device="cuda").requires_grad_(True)
    yt = Top.apply(xt)
    np.testing.assert_array_equal(yt.detach().cpu().numpy(), Op.matmat(x.T).T)
    yt.backward(torch.as_tensor(v, device="cuda"))
    g = xt.grad.cpu().numpy()
    np.testing.assert_array_equal(g, Op.rmatmat(v.T).T)
    for j in range(4):
        assert G.relerr(g[j], orc.rmatvec(v[j].astype(np.float64))) < 1e-5


@pytest.mark.parametrize("case", ["dyn_f32", "dyn_f16", "kin_f32", "kin3d_f16", "neg_byot"])
@pytest.mark.parametrize("k", [2, 8, 13])
def test_matmat_forward_models_bitwise(case, k):
    """The float32 matmat forward spreads 8, 4 or 2 models per workgroup
    (kx_spread_models): every column is bitwise its matvec, for dynamic and
    kinematic operators, float16 tables, 3D, and the negative-time path."""
    import pylops_amd
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav,
              wavcenter=G.wavc, mode="analytic", dtype="float32")
    if case == "dyn_f16":
        kw.update(dynamic=True, tabledtype="float16")
    elif case == "dyn_f32":
        kw.update(dynamic=True)
    elif case == "kin3d_f16":
        kw.update(y=G.y, srcs=G.s3d, recs=G.r3d, tabledtype="float16")
    elif case == "neg_byot":
        ts, tr, _, _, _, _ = pylops_amd.Kirchhoff._traveltime_table(G.z, G.x, G.s2d, G.r2d,
                                                                    G.v0, mode="analytic")
        kw.update(mode="byot", trav=(ts - 0.01113, tr - 0.00871), vel=None)
    Op = pylops_amd.Kirchhoff(**kw)
    rng = np.random.default_rng(k)
    X = rng.standard_normal((Op.shape[1], k)).astype(np.float32)
    X[:, 1] = 0.0  # an all-zero model
    Y = Op.matmat(X)
    for j in range(k):
        np.testing.assert_array_equal(Y[:, j], Op.matvec(X[:, j]))


def test_matmat_forward_models_nonfinite_column():
    """A non-finite model switches its launch to float64 accumulation, as
    kx_spread does: the NaN propagates to that column's traces, the other
    columns stay within rounding of their matvecs."""
    Op = _kop("float32")
    rng = np.random.default_rng(5)
    X = rng.standard_normal((Op.shape[1], 4)).astype(np.float32)
    X[7, 2] = np.nan
    Y = Op.matmat(X)
    for j in range(4):
        want = Op.matvec(X[:, j])
        np.testing.assert_array_equal(np.isnan(Y[:, j]), np.isnan(want))
        ok = ~np.isnan(want)
        assert G.relerr(Y[ok, j], want[ok]) < 1e-6


def test_spread_models_rejects_float64():
    """kx_spread_models is float32-only: a float64 plan is refused (the
    float64 matmat runs the columns through kx_spread)."""
    import ctypes
    import torch
    from pylops_amd import _lib
    Op = _kop("float64")
    m = torch.zeros((2, Op.shape[1]), dtype=torch.float64, device="cuda")
    y = torch.empty((2, Op.shape[0]), dtype=torch.float64, device="cuda")
    rc = _lib.load().kx_spread_models(ctypes.byref(Op.plan), _lib.ptr(m), _lib.ptr(y), 2,
                                      _lib.stream_ptr(torch.device("cuda")))
    assert rc == _lib.KX_ETYPE


@pytest.mark.parametrize("case", ["dyn_f32", "dyn_f16", "kin_f32", "kin3d_f16", "neg_byot"])
@pytest.mark.parametrize("k", [2, 5, 8])
def test_rmatmat_adjoint_models_bitwise(case, k):
    """The float32 rmatmat gathers 4 or 2 models per wave
    (kx_gather_models): every column is bitwise its rmatvec."""
    import pylops_amd
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav,
              wavcenter=G.wavc, mode="analytic", dtype="float32")
    if case == "dyn_f16":
        kw.update(dynamic=True, tabledtype="float16")
    elif case == "dyn_f32":
        kw.update(dynamic=True)
    elif case == "kin3d_f16":
        kw.update(y=G.y, srcs=G.s3d, recs=G.r3d, tabledtype="float16")
    elif case == "neg_byot":
        ts, tr, _, _, _, _ = pylops_amd.Kirchhoff._traveltime_table(G.z, G.x, G.s2d, G.r2d,
                                                                    G.v0, mode="analytic")
        kw.update(mode="byot", trav=(ts - 0.01113, tr - 0.00871), vel=None)
    Op = pylops_amd.Kirchhoff(**kw)
    rng = np.random.default_rng(20 + k)
    V = rng.standard_normal((Op.shape[0], k)).astype(np.float32)
    V[:, 0] = 0.0
    Z = Op.rmatmat(V)
    for j in range(k):
        np.testing.assert_array_equal(Z[:, j], Op.rmatvec(V[:, j]))


@pytest.mark.parametrize("split", [1, 2, 4])
def test_rmatmat_adjoint_models_source_splits(split):
    """Every adjoint source split (waves per tile) keeps the models bitwise
    equal to their own rmatvec."""
    Op = _kop("float32")
    Op.gsplit = split
    rng = np.random.default_rng(31)
    V = rng.standard_normal((Op.shape[0], 6)).astype(np.float32)
    Z = Op.rmatmat(V)
    for j in range(6):
        np.testing.assert_array_equal(Z[:, j], Op.rmatvec(V[:, j]))


def test_matmat_rmatmat_many_columns():
    """More columns than one multi-model call takes (64): the products run
    in chunks and every column stays bitwise its own product."""
    Op = _kop("float32")
    rng = np.random.default_rng(41)
    X = rng.standard_normal((Op.shape[1], 70)).astype(np.float32)
    V = rng.standard_normal((Op.shape[0], 70)).astype(np.float32)
    Y = Op.matmat(X)
    Z = Op.rmatmat(V)
    for j in (0, 7, 63, 64, 69):
        np.testing.assert_array_equal(Y[:, j], Op.matvec(X[:, j]))
        np.testing.assert_array_equal(Z[:, j], Op.rmatvec(V[:, j]))
