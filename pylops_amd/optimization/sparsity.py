"""Function front-ends of ISTA / FISTA (pylops/optimization/sparsity.py:
ista 244-375, fista 378-507): same signatures and return tuples
``(x, niter, cost)``.  As in the reference these take ``y`` as given: a
vector, or a matrix whose columns are several right-hand sides (solved
together through matmat / rmatmat); ``x`` has the matching shape.  numpy
inputs are moved to the operator's device and results come back as numpy;
torch device tensors stay on the device."""
from pylops_amd.optimization.basic import _prepare
from pylops_amd.optimization.cls_sparsity import FISTA, ISTA

__all__ = ["ista", "fista"]


def _run(cls, Op, y, x0, niter, SOp, eps, alpha, eigsdict, tol, threshkind, perc, decay,
         monitorres, show, itershow, callback):
    solver = cls(Op)
    several = y.ndim > 1 and y.shape[1] > 1  # columns are right-hand sides
    yd, x0d, back = _prepare(Op, y, x0, callback, solver, flat=not several)
    x, iiter, cost = solver.solve(yd, x0=x0d, niter=niter, SOp=SOp, eps=eps, alpha=alpha,
                                  eigsdict=eigsdict, tol=tol, threshkind=threshkind, perc=perc,
                                  decay=decay, monitorres=monitorres, show=show,
                                  itershow=itershow)
    return back(x), iiter, cost


def ista(Op, y, x0=None, niter=10, SOp=None, eps=0.1, alpha=None, eigsdict=None, tol=1e-10,
         threshkind="soft", perc=None, decay=None, monitorres=False, show=False,
         itershow=(10, 10, 10), callback=None, preallocate=False):
    """Iterative shrinkage-thresholding; returns ``(x, niter, cost)``."""
    return _run(ISTA, Op, y, x0, niter, SOp, eps, alpha, eigsdict, tol, threshkind, perc,
                decay, monitorres, show, itershow, callback)


def fista(Op, y, x0=None, niter=10, SOp=None, eps=0.1, alpha=None, eigsdict=None, tol=1e-10,
          threshkind="soft", perc=None, decay=None, monitorres=False, show=False,
          itershow=(10, 10, 10), callback=None, preallocate=False):
    """Fast iterative shrinkage-thresholding; returns ``(x, niter, cost)``."""
    return _run(FISTA, Op, y, x0, niter, SOp, eps, alpha, eigsdict, tol, threshkind, perc,
                decay, monitorres, show, itershow, callback)
