"""Device ISTA / FISTA (pylops/optimization/cls_sparsity.py) against the
oracle's restatement (itself pinned by proximal-map, optimality and
recovery checks in tests/test_oracle.py) and the reference's own tests.

Bars: float64 Kirchhoff operator vs the oracle on its dense matrix, same
step size: 1e-9 relative on x and the cost history; reference recovery
tests: one decimal, as the reference."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import geometry as G
from tests.test_solvers_gpu import _TorchDense

pytestmark = pytest.mark.gpu


def test_unknown_threshkind_and_missing_perc():
    """pytests/test_sparsity.py:243-256."""
    from pylops_amd.optimization.sparsity import fista, ista
    Dop = _TorchDense.make(np.eye(5))
    for f in (ista, fista):
        with pytest.raises(NotImplementedError):
            f(Dop, np.ones(5), niter=10, threshkind="foo")
        with pytest.raises(ValueError):
            f(Dop, np.ones(5), niter=10, perc=None, threshkind="soft-percentile")


@pytest.mark.parametrize("shape", [(11, 11), (31, 11), (21, 41)])
def test_ista_fista_recovery(shape):
    """pytests/test_sparsity.py:259-343 on a dense device operator: every
    threshold kind and the percentile kinds recover a 3-spike model to one
    decimal; a step size that is far too large stops on the residual
    monitor."""
    from pylops_amd.optimization.sparsity import fista, ista
    np.random.seed(42)
    A = np.random.randn(*shape)
    Aop = _TorchDense.make(A)
    x = np.zeros(shape[1])
    x[shape[1] // 2] = 1
    x[3] = 1
    x[shape[1] - 4] = -1
    y = A @ x
    with pytest.raises(ValueError):
        ista(Aop, y, niter=2000, eps=0.5, alpha=1e5, monitorres=True, tol=0)
    for kind in ("hard", "soft", "half"):
        for f in (ista, fista):
            xinv, _, _ = f(Aop, y, niter=2000, eps=0.5, threshkind=kind, tol=0)
            np.testing.assert_array_almost_equal(x, xinv, decimal=1)
    for kind in ("hard-percentile", "soft-percentile", "half-percentile"):
        for f in (ista, fista):
            xinv, _, _ = f(Aop, y, niter=2000, perc=30, threshkind=kind, tol=0)
            np.testing.assert_array_almost_equal(x, xinv, decimal=1)


@pytest.mark.parametrize("fast", [True, False])
@pytest.mark.parametrize("kind", ["soft", "half"])
def test_shrinkage_on_kirchhoff_matches_oracle(fast, kind):
    """Device FISTA/ISTA on the float64 HIP Kirchhoff operator vs the
    oracle's iterations on the oracle operator's dense matrix, same alpha."""
    import pylops_amd
    from pylops_amd.optimization.sparsity import fista, ista
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav,
              wavcenter=G.wavc, mode="analytic", dynamic=True, dtype="float64")
    Op = pylops_amd.Kirchhoff(**kw)
    orc = O.KirchhoffOracle(**kw)
    n = Op.shape[1]
    A = np.stack([orc.matvec(e) for e in np.eye(n)], axis=1)
    rng = np.random.default_rng(3)
    m = np.zeros(n)
    m[rng.choice(n, 12, replace=False)] = rng.uniform(-1, 1, 12)
    y = A @ m
    alpha = 1.0 / np.linalg.eigvalsh(A.T @ A).max()
    eps = 1e-2 * float(np.abs(A.T @ y).max())
    f = fista if fast else ista
    x, it, cost = f(Op, y, niter=30, eps=eps, alpha=alpha, threshkind=kind, tol=0)
    xo, ito, costo = O.shrinkage(O.DenseOp(A), y, alpha, niter=30, eps=eps, tol=0,
                                 threshkind=kind, fast=fast)
    assert it == ito == 30
    assert G.relerr(np.asarray(x).ravel(), xo) < 1e-9
    np.testing.assert_allclose(cost, costo, rtol=1e-9)


def test_fista_power_iteration_step():
    """Without alpha the step comes from 10 power iterations of Op^H Op
    (eigs.py:12-80, the reference's GPU path): the estimate is a lower
    bound of the largest eigenvalue within a few percent."""
    import pylops_amd
    from pylops_amd.optimization.cls_sparsity import FISTA
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav,
              wavcenter=G.wavc, mode="analytic", dynamic=False, dtype="float64")
    Op = pylops_amd.Kirchhoff(**kw)
    orc = O.KirchhoffOracle(**kw)
    n = Op.shape[1]
    A = np.stack([orc.matvec(e) for e in np.eye(n)], axis=1)
    lmax = np.linalg.eigvalsh(A.T @ A).max()
    import torch
    s = FISTA(Op)
    s.setup(torch.zeros(Op.shape[0], dtype=torch.float64, device="cuda"), niter=1,
            eigsdict=dict(niter=50, tol=1e-10))
    assert 1.0 / s.alpha <= lmax * (1 + 1e-9)
    assert 1.0 / s.alpha >= 0.95 * lmax


def test_lsm_with_fista():
    """tutorials/lsm.py:125-127: LSM.solve with solver=fista (sparse
    reflectivity): the soft-thresholded model is mostly exact zeros, keeps
    the two reflectors and explains most of the data."""
    from pylops_amd.optimization.sparsity import fista
    from pylops_amd.waveeqprocessing import LSM
    from tests.test_lsm_gpu import r2d, s2d, t, v0, wav, wavc, x, z
    refl = np.zeros((len(x), len(z)))
    refl[:, len(z) // 2] = 1
    refl[3, 3 * len(z) // 4] = -1
    lsm = LSM(z, x, t, s2d, r2d, v0, wav, wavc, mode="analytic")
    d = lsm.Demop @ refl.ravel()
    eps = 0.05 * float(np.abs(lsm.Demop.H @ d).max())
    mf = lsm.solve(d, solver=fista, **dict(eps=eps, niter=200))
    res_f = np.linalg.norm(d - lsm.Demop @ mf.ravel()) / np.linalg.norm(d)
    assert res_f < 0.5
    assert np.mean(mf == 0) > 0.5
    assert np.all(np.abs(mf[:, len(z) // 2]) > 0.3)


@pytest.mark.parametrize("fast", [True, False])
def test_analysis_form_sop_matches_oracle(fast):
    """ISTA/FISTA with SOp (cls_sparsity.py:1734-1758, 2099-2123) on the
    float64 Kirchhoff operator, SOp an orthonormal dense transform on the
    device, vs the oracle's iterations with the same S."""
    import pylops_amd
    from pylops_amd.optimization.sparsity import fista, ista
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav,
              wavcenter=G.wavc, mode="analytic", dynamic=False, dtype="float64")
    Op = pylops_amd.Kirchhoff(**kw)
    orc = O.KirchhoffOracle(**kw)
    n = Op.shape[1]
    A = np.stack([orc.matvec(e) for e in np.eye(n)], axis=1)
    rng = np.random.default_rng(9)
    S, _ = np.linalg.qr(rng.standard_normal((n, n)))
    c = np.zeros(n)
    c[rng.choice(n, 10, replace=False)] = rng.uniform(-1, 1, 10)
    y = A @ (S @ c)
    alpha = 1.0 / np.linalg.eigvalsh(A.T @ A).max()
    eps = 1e-2 * float(np.abs(S.T @ (A.T @ y)).max())
    f = fista if fast else ista
    x, it, cost = f(Op, y, niter=25, SOp=_TorchDense.make(S), eps=eps, alpha=alpha, tol=0)
    xo, ito, costo = O.shrinkage(O.DenseOp(A), y, alpha, niter=25, eps=eps, tol=0,
                                 fast=fast, S=S)
    assert it == ito == 25
    assert G.relerr(np.asarray(x).ravel(), xo) < 1e-9
    np.testing.assert_allclose(cost, costo, rtol=1e-9)


@pytest.mark.parametrize("fast", [True, False])
def test_several_right_hand_sides(fast):
    """y with several columns (cls_sparsity.py:1544-1557, 1633-1646): the
    products go through matmat / rmatmat, every column follows its own
    single right-hand-side iterations (same alpha, elementwise thresholds),
    and the cost uses the matrix norms as numpy's: 0.5 |R|_F^2 + eps |X|_1
    (the largest column l1 norm)."""
    import pylops_amd
    from pylops_amd.optimization.sparsity import fista, ista
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav,
              wavcenter=G.wavc, mode="analytic", dynamic=True, dtype="float64")
    Op = pylops_amd.Kirchhoff(**kw)
    orc = O.KirchhoffOracle(**kw)
    n = Op.shape[1]
    A = np.stack([orc.matvec(e) for e in np.eye(n)], axis=1)
    rng = np.random.default_rng(12)
    M = np.zeros((n, 3))
    for j in range(3):
        M[rng.choice(n, 8, replace=False), j] = rng.uniform(-1, 1, 8)
    Y = A @ M
    alpha = 1.0 / np.linalg.eigvalsh(A.T @ A).max()
    eps = 1e-2 * float(np.abs(A.T @ Y).max())
    f = fista if fast else ista
    X, it, cost = f(Op, Y, niter=20, eps=eps, alpha=alpha, tol=0)
    assert X.shape == (n, 3) and it == 20
    for j in range(3):
        xj = f(Op, Y[:, j], niter=20, eps=eps, alpha=alpha, tol=0)[0]
        np.testing.assert_allclose(X[:, j], xj, rtol=1e-11, atol=1e-14)
    if fast:
        want = 0.5 * np.linalg.norm(Y - A @ X) ** 2 + eps * np.linalg.norm(X, ord=1)
        np.testing.assert_allclose(cost[-1], want, rtol=1e-9)
    with pytest.raises(ValueError):
        f(Op, Y, x0=np.zeros(n), niter=2, eps=eps, alpha=alpha)
