"""ISTA and FISTA with device-resident vectors
(pylops/optimization/cls_sparsity.py: thresholds 36-200, ISTA 1313-1944,
FISTA 1945-2160; step size from pylops/optimization/eigs.py:12-80).

The iterations are those of the reference: the same thresholds, threshold
decay, momentum sequence, stopping rule on the model update, residual
monitor and cost history (FISTA's cost evaluates ``Op @ x`` as the
reference does).  Vectors are torch tensors on the operator's device, the
operator products are the HIP kernels; the step size, when not given, comes
from the power iteration the reference uses on GPU arrays (10 iterations of
``Op^H Op``, ``eigsdict`` passes ``niter`` / ``tol``).  With ``SOp`` the
threshold acts on ``SOp^H`` of the gradient-updated model and ``SOp`` maps
the thresholded coefficients back (cls_sparsity.py:1734-1758, 2099-2123).
"""
import logging
import math
import time

import numpy as np
import torch

from pylops_amd.optimization.basesolver import Solver, _units
from pylops_amd.optimization.cls_basic import _apply_into

__all__ = ["ISTA", "FISTA", "power_iteration"]

logger = logging.getLogger(__name__)

_THRESHKINDS = ("hard", "soft", "half", "hard-percentile", "soft-percentile",
                "half-percentile")


def _hardthreshold(x, thresh):
    x1 = x.clone()
    x1[x.abs() <= math.sqrt(2 * thresh)] = 0
    return x1


def _softthreshold(x, thresh):
    return torch.clamp(x.abs() - thresh, min=0.0) * torch.sign(x)


def _halfthreshold(x, thresh):
    arg = torch.ones_like(x)
    nz = x != 0
    arg[nz] = (thresh / 8.0) * (x[nz].abs() / 3.0) ** (-1.5)
    arg = torch.clamp(arg, -1, 1)
    phi = 2.0 / 3.0 * torch.arccos(arg)
    x1 = 2.0 / 3.0 * x * (1 + torch.cos(2.0 * np.pi / 3.0 - phi))
    x1[x.abs() <= (54 ** (1.0 / 3.0) / 4.0) * thresh ** (2.0 / 3.0)] = 0
    return x1


def _percentile(a, perc):
    """numpy.percentile (linear interpolation) of a device tensor."""
    s = torch.sort(a.reshape(-1))[0]
    pos = perc / 100.0 * (s.numel() - 1)
    lo = int(math.floor(pos))
    hi = min(lo + 1, s.numel() - 1)
    w = pos - lo
    return float(s[lo]) * (1 - w) + float(s[hi]) * w if w else float(s[lo])


def _hardthreshold_percentile(x, perc):
    thresh = _percentile(x.abs(), perc)
    return _hardthreshold(x, 0.5 * thresh ** 2)


def _softthreshold_percentile(x, perc):
    return _softthreshold(x, _percentile(x.abs(), perc))


def _halfthreshold_percentile(x, perc):
    thresh = _percentile(x.abs(), perc)
    return _halfthreshold(x, (4.0 / 54 ** (1.0 / 3.0) * thresh) ** 1.5)


_THRESHF = {"hard": _hardthreshold, "soft": _softthreshold, "half": _halfthreshold,
            "hard-percentile": _hardthreshold_percentile,
            "soft-percentile": _softthreshold_percentile,
            "half-percentile": _halfthreshold_percentile}


def power_iteration(apply, n, niter=10, tol=1e-5, dtype=torch.float32, device=None,
                    generator=None):
    """Largest eigenvalue of the square operator ``apply`` of size ``n``
    (eigs.py:12-80): ``(maxeig, b_k, iiter)``, starting from a uniform
    random vector."""
    b = torch.rand(n, generator=generator, dtype=torch.float64).to(device=device, dtype=dtype)
    b = b / torch.linalg.norm(b)
    niter = 10 if niter is None else niter
    maxeig_old = 0.0
    for iiter in range(niter):
        b1 = apply(b)
        maxeig = float(torch.dot(b.double(), b1.double()))
        b = b1 / torch.linalg.norm(b1)
        if abs(maxeig - maxeig_old) < tol * maxeig_old:
            break
        maxeig_old = maxeig
    return maxeig, b, iiter + 1


class ISTA(Solver):
    r"""Iterative shrinkage-thresholding (cls_sparsity.py:1313-1944):
    minimises :math:`\|\mathbf{y}-\mathbf{Op}\,\mathbf{x}\|_2^2 +
    \epsilon\|\mathbf{x}\|_p` with hard, soft or half thresholding (or their
    percentile forms)."""

    def _print_setup(self):
        self._print_solver(f" (threshkind={self.threshkind})", nbar=80)
        if self.perc is None:
            print(f"eps = {self.eps:10e}\ttol = {self.tol:10e}\tniter = {self.niter}")
            print(f"alpha = {self.alpha:10e}\tthresh = {self.thresh:10e}")
        else:
            print(f"perc = {self.perc:.1f}\ttol = {self.tol:10e}\tniter = {self.niter}")
            print(f"alpha = {self.alpha:10e}")
        print("-" * 80 + "\n")
        print("   Itn          x[0]        r2norm     r12norm     xupdate")

    def _print_step(self, x, costdata, costreg, xupdate):
        print(f"{self.iiter:6g} {float(x.reshape(-1)[0]):12.5e}  {costdata:10.3e}   "
              f"{costdata + costreg:9.3e}  {xupdate:10.3e}")

    def memory_usage(self, show=False, unit="B"):
        nbytes = np.dtype(self.Op.dtype).itemsize
        memuse = (3 * self.Op.shape[1] + 2 * self.Op.shape[0]) * nbytes
        if show:
            print(f"{type(self).__name__} predicted memory usage: "
                  f"{memuse / _units[unit]:.2f} {unit}")
        return memuse

    @staticmethod
    def _apply(Op, x, adjoint):
        """Op x (Op^H x if adjoint) for a vector, or Op.matmat / rmatmat for
        the columns of a matrix (several right-hand sides,
        cls_sparsity.py:1544-1557)."""
        if x.ndim == 2:
            y = Op.rmatmat(x) if adjoint else Op.matmat(x)
            return torch.as_tensor(y, dtype=x.dtype, device=x.device).reshape(
                Op.shape[1] if adjoint else Op.shape[0], x.shape[1])
        out = torch.empty(Op.shape[1] if adjoint else Op.shape[0], dtype=x.dtype,
                          device=x.device)
        return _apply_into(Op, x, out, adjoint)

    def _fwd(self, x):
        return self._apply(self.Op, x, False)

    def _adj(self, r):
        return self._apply(self.Op, r, True)

    def setup(self, y, x0=None, niter=None, SOp=None, eps=0.1, alpha=None, eigsdict=None,
              tol=1e-10, threshkind="soft", perc=None, decay=None, monitorres=False,
              show=False):
        """cls_sparsity.py:1460-1620; returns the initial model."""
        if SOp is not None and (SOp.shape[0] != self.Op.shape[1]):
            raise ValueError(f"SOp of shape {SOp.shape} for a model of {self.Op.shape[1]} "
                             "samples")
        self.SOp = SOp
        if threshkind not in _THRESHKINDS:
            raise NotImplementedError("threshkind should be hard, soft, half,"
                                      "hard-percentile, soft-percentile, "
                                      "or half-percentile")
        if threshkind in _THRESHKINDS[3:] and perc is None:
            raise ValueError("Provide a percentile when choosing hard-percentile,"
                             "soft-percentile, or half-percentile thresholding")
        self.y = y
        self.niter = niter
        self.eps = eps
        self.eigsdict = {} if eigsdict is None else dict(eigsdict)
        self.tol = tol
        self.threshkind = threshkind
        self.perc = perc
        self.monitorres = monitorres
        self.threshf = _THRESHF[threshkind]
        dev, dt = y.device, y.dtype
        # threshold decay per iteration (ones unless given)
        self.decay = None if decay is None else np.asarray(decay)
        if alpha is not None:
            self.alpha = alpha
        elif not hasattr(self, "alpha"):
            maxeig = abs(power_iteration(lambda v: self._adj(self._fwd(v)), self.Op.shape[1],
                                         dtype=dt, device=dev, **self.eigsdict)[0])
            self.alpha = 1.0 / maxeig
        self.thresh = eps * self.alpha * 0.5
        if x0 is None:
            shape = (self.Op.shape[1],) if y.ndim == 1 else (self.Op.shape[1], y.shape[1])
            x = torch.zeros(shape, dtype=dt, device=dev)
        else:
            if y.ndim != x0.ndim:
                raise ValueError("Number of columns of x0 and data are not the same")
            if x0.shape[0] != self.Op.shape[1]:
                raise ValueError("Operator and input vector have different dimensions")
            x = x0.to(device=dev, dtype=dt).clone()
        if monitorres:
            self.normresold = np.inf
        self.t = 1.0
        self.cost = []
        self.iiter = 0
        if show:
            self._print_setup()
        return x

    def _threshold(self, xu):
        """Threshold the gradient-updated model, or its SOp^H projection
        followed by SOp (cls_sparsity.py:1734-1758)."""
        if self.SOp is not None:
            xu = self._apply(self.SOp, xu, True)
        if self.perc is None:
            d = 1.0 if self.decay is None else float(self.decay[self.iiter])
            x = self.threshf(xu, d * self.thresh)
        else:
            x = self.threshf(xu, 100 - self.perc)
        if self.SOp is not None:
            x = self._apply(self.SOp, x, False)
        return x

    def _monitor(self, res):
        if self.monitorres:
            self.normres = float(torch.linalg.norm(res))
            if self.normres > self.normresold:
                raise ValueError(f"{type(self).__name__} stopped at iteration {self.iiter} "
                                 "due to residual increasing, consider modifying "
                                 "eps and/or alpha...")
            self.normresold = self.normres

    def step(self, x, show=False):
        """One ISTA iteration (cls_sparsity.py:1700-1800)."""
        xold = x.clone()
        res = self.y - self._fwd(x)
        self._monitor(res)
        grad = self.alpha * self._adj(res)
        x = self._threshold(x + grad)
        xupdate = float(torch.linalg.norm(x - xold))
        costdata = 0.5 * float(torch.linalg.norm(res)) ** 2
        costreg = self.eps * float(torch.linalg.norm(x, ord=1))
        self.cost.append(float(costdata + costreg))
        self.iiter += 1
        if show:
            self._print_step(x, costdata, costreg, xupdate)
        return x, xupdate

    def _showstep(self, show, niter, itershow):
        return show and (self.iiter < itershow[0] or niter - self.iiter < itershow[1]
                         or self.iiter % itershow[2] == 0)

    def run(self, x, niter=None, show=False, itershow=(10, 10, 10)):
        niter = self.niter if niter is None else niter
        if niter is None:
            raise ValueError("niter must not be None")
        xupdate = np.inf
        while self.iiter < niter and xupdate > self.tol:
            x, xupdate = self.step(x, self._showstep(show, niter, itershow))
            self.callback(x)
        if xupdate <= self.tol:
            logger.warning("Update smaller that tolerance for iteration %d", self.iiter)
        return x

    def finalize(self, show=False):
        self.tend = time.time()
        self.telapsed = self.tend - self.tstart
        self.cost = np.array(self.cost)
        if show:
            self._print_finalize()

    def solve(self, y, x0=None, niter=None, SOp=None, eps=0.1, alpha=None, eigsdict=None,
              tol=1e-10, threshkind="soft", perc=None, decay=None, monitorres=False, show=False,
              itershow=(10, 10, 10)):
        x = self.setup(y=y, x0=x0, niter=niter, SOp=SOp, eps=eps, alpha=alpha,
                       eigsdict=eigsdict, tol=tol, threshkind=threshkind, perc=perc, decay=decay,
                       monitorres=monitorres, show=show)
        x = self.run(x, niter, show=show, itershow=itershow)
        self.finalize(show)
        return x, self.iiter, self.cost


class FISTA(ISTA):
    r"""Fast iterative shrinkage-thresholding (cls_sparsity.py:1945-2160):
    ISTA on an auxiliary point with the Beck-Teboulle momentum."""

    def memory_usage(self, show=False, unit="B"):
        nbytes = np.dtype(self.Op.dtype).itemsize
        memuse = (5 * self.Op.shape[1] + 2 * self.Op.shape[0]) * nbytes
        if show:
            print(f"FISTA predicted memory usage: {memuse / _units[unit]:.2f} {unit}")
        return memuse

    def step(self, x, z, show=False):
        """One FISTA iteration (cls_sparsity.py:2010-2120)."""
        xold = x.clone()
        res = self.y - self._fwd(z)
        self._monitor(res)
        grad = self.alpha * self._adj(res)
        x = self._threshold(z + grad)
        told = self.t
        self.t = (1.0 + math.sqrt(1.0 + 4.0 * self.t ** 2)) / 2.0
        z = x + ((told - 1.0) / self.t) * (x - xold)
        xupdate = float(torch.linalg.norm(x - xold))
        costdata = 0.5 * float(torch.linalg.norm(self.y - self._fwd(x))) ** 2
        costreg = self.eps * float(torch.linalg.norm(x, ord=1))
        self.cost.append(float(costdata + costreg))
        self.iiter += 1
        if show:
            self._print_step(x, costdata, costreg, xupdate)
        return x, z, xupdate

    def run(self, x, niter=None, show=False, itershow=(10, 10, 10)):
        niter = self.niter if niter is None else niter
        if niter is None:
            raise ValueError("niter must not be None")
        z = x.clone()
        xupdate = np.inf
        while self.iiter < niter and xupdate > self.tol:
            x, z, xupdate = self.step(x, z, self._showstep(show, niter, itershow))
            self.callback(x)
        if xupdate <= self.tol:
            logger.warning("Update smaller that tolerance for iteration %d", self.iiter)
        return x
