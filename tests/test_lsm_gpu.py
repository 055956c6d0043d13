"""LSM application (pytests/test_lsm.py:60-113 restated on the device
solvers), analytic and eikonal tables (the eikonal ones from the device
eikonal solver; the reference needs scikit-fmm for them)."""
import numpy as np
import pytest
from numpy.testing import assert_array_almost_equal

from oracle import oracle as O

pytestmark = pytest.mark.gpu

PAR = {"nx": 12, "nz": 20, "nt": 50, "dx": 1, "dz": 2, "dt": 0.004, "nsx": 6, "nrx": 4}
v0 = 500
x = np.arange(PAR["nx"]) * PAR["dx"]
z = np.arange(PAR["nz"]) * PAR["dz"]
t = np.arange(PAR["nt"]) * PAR["dt"]
sx = np.linspace(x.min(), x.max(), PAR["nsx"])
rx = np.linspace(x.min(), x.max(), PAR["nrx"])
s2d = np.vstack((sx, 2 * np.ones(PAR["nsx"])))
r2d = np.vstack((rx, 2 * np.ones(PAR["nrx"])))
wav, _, wavc = O.ricker(t[:21], f0=40)


def test_unknown_mode():
    from pylops_amd.waveeqprocessing import LSM
    with pytest.raises(NotImplementedError):
        LSM(z, x, t, s2d, r2d, 0, np.ones(3), 1, mode="foo")
    with pytest.raises(NotImplementedError):
        LSM(z, x, t, s2d, r2d, v0, wav, wavc, kind="twowayac")


@pytest.mark.parametrize("mode", ["analytic", "eikonal"])
@pytest.mark.parametrize("dynamic", [False, True])
def test_lsm2d(dynamic, mode):
    """Dot-test and inversion (test_lsm.py:74-113, par1/par2/par1d/par2d):
    data fit to 2 decimals, reflectivity to 1, with the scipy-lsqr-style
    default solver."""
    from pylops_amd.waveeqprocessing import LSM
    vel = v0 * np.ones((PAR["nx"], PAR["nz"]))
    refl = np.zeros((PAR["nx"], PAR["nz"]))
    refl[:, PAR["nz"] // 2] = 1
    refl[:, 3 * PAR["nz"] // 4] = 1
    lsm = LSM(z, x, t, s2d, r2d, vel if mode == "eikonal" else v0, wav, wavc, mode=mode,
              dynamic=dynamic, dottest=True)
    d = lsm.Demop * refl.ravel()
    d = d.reshape(PAR["nsx"], PAR["nrx"], PAR["nt"])
    minv = lsm.solve(d.ravel(), **dict(iter_lim=100, show=True))
    minv = minv.reshape(PAR["nx"], PAR["nz"])
    dinv = lsm.Demop * minv.ravel()
    dinv = dinv.reshape(PAR["nsx"], PAR["nrx"], PAR["nt"])
    assert_array_almost_equal(d / d.max(), dinv / d.max(), decimal=2)
    assert_array_almost_equal(refl / refl.max(), minv / refl.max(), decimal=1)


def test_lsm_scipy_lsqr_runs_on_the_device():
    """solver=scipy.sparse.linalg.lsqr is routed to the device LSQR with
    scipy's argument names; the result matches scipy's own lsqr on the
    oracle's dense matrix (same recurrences, float64)."""
    import scipy.sparse.linalg as spla
    from pylops_amd.waveeqprocessing import LSM
    lsm = LSM(z, x, t, s2d, r2d, v0, wav, wavc, mode="analytic", dtype="float64")
    rng = np.random.default_rng(0)
    d = rng.standard_normal(lsm.Demop.shape[0])
    m = lsm.solve(d, solver=spla.lsqr, iter_lim=15, atol=0, btol=0)
    orc = O.KirchhoffOracle(z=z, x=x, t=t, srcs=s2d, recs=r2d, vel=v0, wav=wav,
                            wavcenter=wavc, mode="analytic", dtype="float64")
    n = lsm.Demop.shape[1]
    A = np.stack([orc.matvec(e) for e in np.eye(n)], axis=1)
    mref = spla.lsqr(A, d, iter_lim=15, atol=0, btol=0)[0]
    assert np.linalg.norm(m.ravel() - mref) <= 1e-8 * np.linalg.norm(mref)
