"""Least-squares migration (pylops/waveeqprocessing/lsm.py:13-174).

``LSM`` builds the Kirchhoff demigration operator and inverts it with a
chosen solver.  The reference's default solver is
``scipy.sparse.linalg.lsqr``, which would run its iterations on the host
and call the operator with host arrays; here that default (and scipy's
``lsqr`` passed explicitly) runs this package's device LSQR with scipy's
argument names and defaults (``iter_lim = 2 n``, ``atol = btol = 1e-6``,
``calc_var=False``) and returns scipy's 10-tuple.  Any other solver is
called as the reference calls it: ``solver(Demop, d.ravel(), **kwargs)``.
"""
import numpy as np

from pylops_amd.utils.dottest import dottest as _dottest
from pylops_amd.waveeqprocessing.kirchhoff import Kirchhoff

__all__ = ["LSM", "lsqr_scipy"]


def lsqr_scipy(A, b, damp=0.0, atol=1e-6, btol=1e-6, conlim=1e8, iter_lim=None, show=False,
               calc_var=False, x0=None):
    """``scipy.sparse.linalg.lsqr`` signature and return tuple ``(x, istop,
    itn, r1norm, r2norm, anorm, acond, arnorm, xnorm, var)``, computed by
    the device LSQR (the same Paige-Saunders recurrences) with its
    iterations replayed from a HIP graph."""
    from pylops_amd.optimization.basic import lsqr
    niter = 2 * A.shape[1] if iter_lim is None else int(iter_lim)
    (x, istop, itn, r1norm, r2norm, anorm, acond, arnorm, xnorm, var,
     _) = lsqr(A, b, x0=x0, damp=damp, atol=atol, btol=btol, conlim=conlim, niter=niter,
               calc_var=calc_var, show=show, graph=not show)
    if var is None:
        var = np.zeros(A.shape[1])
    return x, istop, itn, r1norm, r2norm, anorm, acond, arnorm, xnorm, var


def _device_solver(solver):
    try:
        import scipy.sparse.linalg as spla
        if solver is spla.lsqr:
            return lsqr_scipy
    except ImportError:  # pragma: no cover
        pass
    return solver


class LSM:
    """Least-squares migration (lsm.py:13-174).

    Parameters are the reference's: ``z, x, t, srcs, recs, vel, wav,
    wavcenter, y=None, kind="kirchhoff", dottest=False, **kwargs_mod``
    (``kwargs_mod`` go to :class:`Kirchhoff`).  ``kind="twowayac"`` (the
    two-way acoustic wave-equation operator) is not provided."""

    def __init__(self, z, x, t, srcs, recs, vel, wav, wavcenter, y=None, kind="kirchhoff",
                 dottest=False, **kwargs_mod):
        self.y, self.x, self.z = y, x, z
        if kind == "kirchhoff":
            self.Demop = Kirchhoff(z, x, t, srcs, recs, vel, wav, wavcenter, y=y, **kwargs_mod)
        elif kind == "twowayac":
            raise NotImplementedError("kind='twowayac' (AcousticWave2D) is not provided by "
                                      "pylops_amd")
        else:
            raise NotImplementedError("kind must be kirchhoff or twowayac")
        if dottest:
            _dottest(self.Demop, self.Demop.shape[0], self.Demop.shape[1], raiseerror=True,
                     verb=True)

    def solve(self, d, solver=lsqr_scipy, **kwargs_solver):
        """Invert ``d`` (ns x nr x nt, or flat) with ``solver``; returns the
        reflectivity of shape (nx, nz) or (ny, nx, nz)."""
        solver = _device_solver(solver)
        minv = solver(self.Demop, d.ravel(), **kwargs_solver)[0]
        if self.y is None:
            return minv.reshape(len(self.x), len(self.z))
        return minv.reshape(len(self.y), len(self.x), len(self.z))
