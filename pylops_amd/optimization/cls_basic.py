"""CGLS and LSQR with device-resident vectors and scalars
(pylops/optimization/cls_basic.py: CGLS CB:334-708, LSQR CB:711-1351).

The vectors are torch tensors on the HIP device.  Every inner product and
norm is accumulated in float64 by libkirchx, and the scalar updates that
the reference performs on the host run as single-thread device kernels, so
a solve enqueues its iterations without waiting on the device.  The loop
exit of the reference becomes a device flag: iterations past it leave every
vector and scalar untouched, so the results equal the reference's early
stop.  With a callback, ``show=True`` or a callbacks list the loop
synchronises every iteration and evaluates the reference's condition on
the host.
"""
import ctypes
import gc
import time

import numpy as np
import torch

from pylops_amd import _lib
from pylops_amd.optimization.basesolver import Solver, _units

__all__ = ["CGLS", "LSQR"]

# how many iterations are enqueued between two looks at the loop flag
_CHECK_EVERY = 8


class _Dev:
    """Launch helpers bound to one device, vector dtype and scalar array."""

    def __init__(self, device, dtype, nsc):
        self.L = _lib.load()
        self.device = device
        self.dtype = dtype
        self.kx = _lib.KX_F64 if dtype == torch.float64 else _lib.KX_F32
        self.sc = torch.zeros(nsc, dtype=torch.float64, device=device)
        nw = (self.L.kx_solver_workspace_bytes() + 7) // 8
        self.work = torch.zeros(nw, dtype=torch.float64, device=device)
        self.base = self.sc.data_ptr()

    def s(self, i):
        """Device address of sc[i]."""
        return ctypes.c_void_p(self.base + 8 * i)

    def stream(self):
        return _lib.stream_ptr(self.device)

    def vec(self, op, x, y, coef, sign=1.0, out=None, gate0=None, gate1=None):
        rc = self.L.kx_vec(op, y.numel(), self.kx, _lib.ptr(x), _lib.ptr(y),
                           self.s(coef) if coef is not None else None, float(sign),
                           self.s(out) if out is not None else None, _lib.ptr(self.work),
                           self.s(gate0) if gate0 is not None else None,
                           self.s(gate1) if gate1 is not None else None, self.stream())
        _lib.check(rc, "kx_vec")

    def dot(self, a, b, out, gate0=None):
        rc = self.L.kx_dot_dev(a.numel(), self.kx, _lib.ptr(a), _lib.ptr(b), self.s(out),
                               _lib.ptr(self.work), self.s(gate0) if gate0 is not None else None,
                               self.stream())
        _lib.check(rc, "kx_dot_dev")

    def host(self):
        return self.sc.cpu().numpy()


def _flat(v, device, dtype):
    return v.reshape(-1).to(device=device, dtype=dtype)


def _apply_into(Op, x, out, adjoint):
    """out <- Op x (Op^H x if adjoint).  Operators with a device fast path
    (the Kirchhoff operator) write straight into ``out``; others go through
    matvec/rmatvec and a copy."""
    into = getattr(Op, "_into", None)
    if into is not None and torch.empty(0, dtype=out.dtype).numpy().dtype == np.dtype(Op.dtype):
        return into(x, out, adjoint)
    y = Op.rmatvec(x) if adjoint else Op.matvec(x)
    out.copy_(_flat(y, out.device, out.dtype))
    return out


def _replay(graph, n, sc, active_index, device):
    """Replay a captured iteration n times, looking at the loop flag every
    few replays so that a converged solve stops early.  Python's cyclic
    garbage collector is held off meanwhile: the loop allocates next to
    nothing, and a full collection (tens of ms in a process with many
    objects) just after a flag check would leave the device idle."""
    enabled = gc.isenabled()
    gc.disable()
    try:
        for k in range(n):
            if k % _CHECK_EVERY == 0 and k > 0 and sc[active_index].item() == 0.0:
                break
            graph.replay()
    finally:
        if enabled:
            gc.enable()


_capture_streams = {}


def _recount(Op, counts, iters):
    """Set the operator's product counters to their values at ``counts`` plus
    one matvec and one rmatvec per iteration run: a replayed graph never
    passes through the counting wrappers, and iterations enqueued after
    convergence are skipped on the device (CGLS.step and LSQR.step of
    CB:510-564, 987-1136 each apply Op and Op^H once)."""
    if counts is not None:
        Op.matvec_count = counts[0] + iters
        Op.rmatvec_count = counts[1] + iters


def _counts(Op):
    if hasattr(Op, "matvec_count") and hasattr(Op, "rmatvec_count"):
        return Op.matvec_count, Op.rmatvec_count
    return None


def _capture(body, device):
    """Capture ``body`` in a HIP graph on a side stream (one per device,
    default priority, reused).  torch's ``graph`` context would first empty
    the allocator's cache, so that every solve allocated its device vectors
    afresh; capture_begin / capture_end keep the cache.  Python's cyclic
    garbage collector is kept off during the capture: collecting an earlier
    solver there would destroy its graph while the stream is capturing,
    which HIP forbids."""
    graph = torch.cuda.CUDAGraph()
    torch.cuda.synchronize(device)
    dev = torch.device(device)
    key = dev.index if dev.index is not None else torch.cuda.current_device()
    stream = _capture_streams.get(key)
    if stream is None:
        stream = _capture_streams[key] = torch.cuda.Stream(device=device)
    enabled = gc.isenabled()
    gc.disable()
    try:
        with torch.cuda.stream(stream):
            graph.capture_begin()
            try:
                body()
            finally:
                graph.capture_end()
    finally:
        if enabled:
            gc.enable()
    return graph


class CGLS(Solver):
    r"""Conjugate gradient least squares: minimises
    :math:`\|\mathbf{y} - \mathbf{Op}\,\mathbf{x}\|_2^2 + \epsilon^2\|\mathbf{x}\|_2^2`
    (CB:334-708).  Vectors are device tensors; inner products in float64."""

    def _print_setup(self, xcomplex=False):
        self._print_solver(nbar=65)
        if self.niter is not None:
            strpar = f"damp = {self.damp:10e}\ttol = {self.tol:10e}\tniter = {self.niter}"
        else:
            strpar = f"damp = {self.damp:10e}\ttol = {self.tol:10e}\t"
        print(strpar)
        print("-" * 65 + "\n")
        print("    Itn          x[0]              r1norm         r2norm")

    def _print_step(self, x):
        c = self._cost_host()
        print(f"{self.iiter:6g}       {float(x[0]):11.4e}        "
              f"{c[0][self.iiter]:11.4e}    {c[1][self.iiter]:11.4e}")

    def memory_usage(self, show=False, unit="B"):
        nbytes = np.dtype(self.Op.dtype).itemsize
        memuse = (2 * self.Op.shape[1] + 3 * self.Op.shape[0]) * nbytes
        memuse += (3 * self.Op.shape[1]) * nbytes
        if show:
            print(f"CGLS predicted memory usage: {memuse / _units[unit]:.2f} {unit}")
        return memuse

    def _cost_host(self):
        n = self.iiter + 1
        return self._cost[:n].cpu().numpy(), self._cost1[:n].cpu().numpy()

    def _grow(self, need):
        if need > self._cost.numel():
            k = max(need, 2 * self._cost.numel())
            for name in ("_cost", "_cost1"):
                old = getattr(self, name)
                new = torch.zeros(k, dtype=torch.float64, device=old.device)
                new[:old.numel()] = old
                setattr(self, name, new)

    def setup(self, y, x0=None, niter=None, damp=0.0, tol=1e-4, preallocate=False,
              show=False):
        """CB:418-508.  ``y`` is a device tensor; returns the initial x."""
        Op = self.Op
        self.y = y
        self.damp = damp ** 2
        self.tol = tol
        self.niter = niter
        self.preallocate = preallocate
        dev, dt = y.device, y.dtype
        D = self._d = _Dev(dev, dt, _lib.KX_CG_NSC)
        sc = D.sc
        if x0 is None:
            x = torch.zeros(Op.shape[1], dtype=dt, device=dev)
            self.s = y.clone()
            self.c = _flat(Op.rmatvec(self.s), dev, dt).clone()
        else:
            x = _flat(x0, dev, dt).clone()
            self.s = y - _flat(Op.matvec(x), dev, dt)
            self.c = _flat(Op.rmatvec(self.s), dev, dt) - damp * x  # raw damp (CB:477)
        self.q = torch.empty(Op.shape[0], dtype=dt, device=dev)
        self.r = torch.empty(Op.shape[1], dtype=dt, device=dev)
        _apply_into(Op, self.c, self.q, False)
        D.dot(self.c, self.c, _lib.KX_CG_KOLD)
        D.dot(self.c, self.c, _lib.KX_CG_CC)
        D.dot(self.q, self.q, _lib.KX_CG_QQ)
        D.dot(self.s, self.s, _lib.KX_CG_SS)
        D.dot(x, x, _lib.KX_CG_XX)
        n0 = (niter if niter is not None else 64) + 1
        self._cost = torch.zeros(n0, dtype=torch.float64, device=dev)
        self._cost1 = torch.zeros(n0, dtype=torch.float64, device=dev)
        self._cost[0] = torch.sqrt(sc[_lib.KX_CG_SS])
        self._cost1[0] = torch.sqrt(sc[_lib.KX_CG_SS] + damp * sc[_lib.KX_CG_XX].abs())
        sc[_lib.KX_CG_DAMP2] = self.damp
        sc[_lib.KX_CG_TOL] = tol
        sc[_lib.KX_CG_ITER] = 0.0
        sc[_lib.KX_CG_NITER] = float(niter) if niter is not None else float("inf")
        sc[_lib.KX_CG_ACTIVE] = (sc[_lib.KX_CG_KOLD] > tol).double()
        self.iiter = 0
        if show:
            self._print_setup()
        return x

    def _step_body(self, x):
        """Enqueue one CGLS iteration (CB:510-564); touches only persistent
        buffers, so it can be captured in a graph and replayed."""
        D = self._d
        L = D.L
        A = _lib.KX_CG_ACTIVE
        st = D.stream()
        cost, cost1 = _lib.ptr(self._cost), _lib.ptr(self._cost1)
        _lib.check(L.kx_cgls_scalar(D.s(0), cost, cost1, 0, st), "kx_cgls_scalar")
        D.vec(_lib.KX_VEC_AXPY, self.c, x, _lib.KX_CG_A, 1.0, _lib.KX_CG_XX, A)
        D.vec(_lib.KX_VEC_AXPY, self.q, self.s, _lib.KX_CG_A, -1.0, _lib.KX_CG_SS, A)
        _apply_into(self.Op, self.s, self.r, True)
        if self.damp != 0:
            D.vec(_lib.KX_VEC_AXPY, x, self.r, _lib.KX_CG_DAMP2, -1.0, _lib.KX_CG_K, A)
        else:
            D.dot(self.r, self.r, _lib.KX_CG_K, A)
        _lib.check(L.kx_cgls_scalar(D.s(0), cost, cost1, 1, st), "kx_cgls_scalar")
        D.vec(_lib.KX_VEC_XPBY, self.r, self.c, _lib.KX_CG_B, 1.0, _lib.KX_CG_CC, A)
        _apply_into(self.Op, self.c, self.q, False)
        D.dot(self.q, self.q, _lib.KX_CG_QQ, A)
        _lib.check(L.kx_cgls_scalar(D.s(0), cost, cost1, 2, st), "kx_cgls_scalar")

    def step(self, x, show=False):
        """One CGLS iteration (CB:510-564), enqueued on the device."""
        self._grow(self.iiter + 2)
        self._step_body(x)
        self.iiter += 1
        if show:
            self._print_step(x)
        return x

    def _sync_state(self):
        h = self._d.host()
        self.iiter = int(h[_lib.KX_CG_ITER])
        return h

    def run(self, x, niter=None, show=False, itershow=(10, 10, 10), graph=False):
        """CB:566-611.  ``graph=True`` runs the first iteration eagerly,
        captures the next one in a HIP graph and replays it."""
        niter = self.niter if niter is None else niter
        if niter is None:
            raise ValueError("niter must not be None")
        counts, it0 = _counts(self.Op), self.iiter
        self._grow(niter + 1)
        self._d.sc[_lib.KX_CG_NITER] = float(niter)
        self._d.sc[_lib.KX_CG_ACTIVE] = (
            (self._d.sc[_lib.KX_CG_ITER] < niter)
            & (self._d.sc[_lib.KX_CG_KOLD] > self.tol)).double()
        if show or self._wants_host_each_step():
            while True:
                h = self._sync_state()
                if not (self.iiter < niter and abs(h[_lib.KX_CG_KOLD]) > self.tol):
                    break
                showstep = show and (self.iiter < itershow[0] or niter - self.iiter < itershow[1]
                                     or self.iiter % itershow[2] == 0)
                x = self.step(x, showstep)
                self.callback(x)
        elif graph and niter - self.iiter > 1:
            x = self.step(x)
            g = _capture(lambda: self._step_body(x), self._d.device)
            _replay(g, niter - self.iiter, self._d.sc, _lib.KX_CG_ACTIVE, self._d.device)
            self._graph = g
            self._sync_state()
        else:
            start = self.iiter
            for k in range(max(niter - start, 0)):
                if k % _CHECK_EVERY == 0 and k > 0:
                    if self._d.sc[_lib.KX_CG_ACTIVE].item() == 0.0:
                        break
                x = self.step(x)
            self._sync_state()
        _recount(self.Op, counts, self.iiter - it0)
        return x

    def finalize(self, show=False):
        """CB:613-630: istop from kold vs tol, r1norm = kold (the reference
        returns the squared normal-equation residual), r2norm = cost1."""
        self.tend = time.time()
        self.telapsed = self.tend - self.tstart
        h = self._sync_state()
        kold = float(abs(h[_lib.KX_CG_KOLD]))
        self.kold = kold
        self.istop = 1 if kold < self.tol else 2
        self.r1norm = kold
        cost, cost1 = self._cost_host()
        self.r2norm = float(cost1[self.iiter])
        self.cost = cost
        self.cost1 = cost1
        if show:
            self._print_finalize(nbar=65)

    def solve(self, y, x0=None, niter=10, damp=0.0, tol=1e-4, preallocate=False, show=False,
              itershow=(10, 10, 10), graph=False):
        x = self.setup(y=y, x0=x0, niter=niter, damp=damp, tol=tol, preallocate=preallocate,
                       show=show)
        x = self.run(x, niter, show=show, itershow=itershow, graph=graph)
        self.finalize(show)
        return x, self.istop, self.iiter, self.r1norm, self.r2norm, self.cost


class LSQR(Solver):
    r"""LSQR (Paige and Saunders, 1982) as CB:711-1351, with device vectors,
    float64 norms and the scalar recurrences on the device."""

    def __init__(self, Op, callbacks=None):
        super().__init__(Op, callbacks)
        self.msg = (
            "The exact solution is x = 0                               ",
            "Op x - b is small enough, given atol, btol                 ",
            "The least-squares solution is good enough, given atol     ",
            "The estimate of cond(Opbar) has exceeded conlim            ",
            "Op x - b is small enough for this machine                  ",
            "The least-squares solution is good enough for this machine",
            "Cond(Opbar) seems to be too large for this machine         ",
            "The iteration limit has been reached                      ",
        )

    def _print_setup(self, x, xcomplex=False):
        self._print_solver(nbar=90)
        print(f"damp = {self.damp:20.14e}     calc_var = {self.calc_var:6g}")
        print(f"atol = {self.atol:8.2e}                 conlim = {self.conlim:8.2e}")
        if self.niter is not None:
            print(f"btol = {self.btol:8.2e}                 niter = {self.niter:8g}")
        else:
            print(f"btol = {self.btol:8.2e}")
        print("-" * 90)
        print("   Itn     x[0]      r1norm     r2norm   Compatible   LS     Norm A   Cond A")
        test2 = self.alfa / self.beta if self.beta else float("nan")
        print(f"{0:6g} {float(x[0]):11.4e} {self.r1norm:10.3e} {self.r2norm:10.3e}"
              f"  {1:8.1e} {test2:8.1e}")

    def _print_step(self, x):
        h = self._d.host()
        print(f"{self.iiter:6g} {float(x[0]):11.4e} {h[_lib.KX_LS_R1NORM]:10.3e} "
              f"{h[_lib.KX_LS_R2NORM]:10.3e}  {h[_lib.KX_LS_TEST1]:8.1e} "
              f"{h[_lib.KX_LS_TEST2]:8.1e} {h[_lib.KX_LS_ANORM]:8.1e} {h[_lib.KX_LS_ACOND]:8.1e}")

    def _print_finalize(self):
        print(" ")
        print(f"LSQR finished, {self.msg[self.istop]}")
        print(" ")
        print(f"istop ={self.istop:8g}   r1norm ={self.r1norm:8.1e}   "
              f"anorm ={self.anorm:8.1e}   arnorm ={self.arnorm:8.1e}")
        print(f"itn   ={self.iiter:8g}   r2norm ={self.r2norm:8.1e}   "
              f"acond ={self.acond:8.1e}   xnorm  ={self.xnorm:8.1e}")
        print(f"Total time (s) = {self.telapsed:.2f}")
        print("-" * 90 + "\n")

    def memory_usage(self, show=False, unit="B"):
        nbytes = np.dtype(self.Op.dtype).itemsize
        memuse = (4 * self.Op.shape[1] + 2 * self.Op.shape[0]) * nbytes
        memuse += self.Op.shape[1] * nbytes
        if show:
            print(f"LSQR predicted memory usage: {memuse / _units[unit]:.2f} {unit}")
        return memuse

    def _grow(self, need):
        if need > self._cost.numel():
            k = max(need, 2 * self._cost.numel())
            new = torch.zeros(k, dtype=torch.float64, device=self._cost.device)
            new[:self._cost.numel()] = self._cost
            self._cost = new

    def setup(self, y, x0=None, damp=0.0, atol=1e-08, btol=1e-08, conlim=100000000.0,
              niter=10, calc_var=True, preallocate=False, show=False):
        """CB:835-985 (host branches on beta and alfa, evaluated once)."""
        Op = self.Op
        self.y = y
        self.damp = damp
        self.atol = atol
        self.btol = btol
        self.conlim = conlim
        self.niter = niter
        self.calc_var = calc_var
        self.preallocate = preallocate
        dev, dt = y.device, y.dtype
        D = self._d = _Dev(dev, dt, _lib.KX_LS_NSC)
        m, n = Op.shape
        self.iiter = 0
        self.istop = 0
        ctol = 1.0 / conlim if conlim > 0 else 0.0
        if x0 is None:
            x = torch.zeros(n, dtype=dt, device=dev)
            self.u = y.clone()
        else:
            x = _flat(x0, dev, dt).clone()
            self.u = y - _flat(Op.matvec(x), dev, dt)
        D.dot(self.u, self.u, _lib.KX_LS_UU)
        beta = float(np.sqrt(D.sc[_lib.KX_LS_UU].item()))
        alfa = 0.0
        if beta > 0.0:
            D.sc[_lib.KX_LS_BETA] = beta
            D.vec(_lib.KX_VEC_SCALE, None, self.u, _lib.KX_LS_BETA)
            self.v = _flat(Op.rmatvec(self.u), dev, dt).clone()
            D.dot(self.v, self.v, _lib.KX_LS_VV)
            alfa = float(np.sqrt(D.sc[_lib.KX_LS_VV].item()))
            if alfa > 0:
                D.sc[_lib.KX_LS_ALFA] = alfa
                D.vec(_lib.KX_VEC_SCALE, None, self.v, _lib.KX_LS_ALFA)
        else:
            self.v = x.clone()
            alfa = 0.0
        self.w = self.v.clone()
        self.tmp_d = torch.empty(m, dtype=dt, device=dev)
        self.tmp_m = torch.empty(n, dtype=dt, device=dev)
        self.alfa, self.beta = alfa, beta
        arnorm = alfa * beta
        self.arnorm = arnorm
        self.arnorm0 = arnorm
        self.rnorm = beta
        self.r1norm = beta
        self.r2norm = beta
        vals = {
            _lib.KX_LS_ALFA: alfa, _lib.KX_LS_BETA: beta, _lib.KX_LS_ANORM: 0.0,
            _lib.KX_LS_ACOND: 0.0, _lib.KX_LS_DAMP: damp, _lib.KX_LS_DAMPSQ: damp ** 2,
            _lib.KX_LS_DDNORM: 0.0, _lib.KX_LS_RES2: 0.0, _lib.KX_LS_XNORM: 0.0,
            _lib.KX_LS_XXNORM: 0.0, _lib.KX_LS_Z: 0.0, _lib.KX_LS_CS2: -1.0,
            _lib.KX_LS_SN2: 0.0, _lib.KX_LS_RHOBAR: alfa, _lib.KX_LS_PHIBAR: beta,
            _lib.KX_LS_BNORM: beta, _lib.KX_LS_RNORM: beta, _lib.KX_LS_R1NORM: beta,
            _lib.KX_LS_R2NORM: beta, _lib.KX_LS_ARNORM: arnorm, _lib.KX_LS_ARNORM0: arnorm,
            _lib.KX_LS_ATOL: atol, _lib.KX_LS_BTOL: btol, _lib.KX_LS_CTOL: ctol,
            _lib.KX_LS_ISTOP: 0.0, _lib.KX_LS_ITER: 0.0,
            _lib.KX_LS_NITER: float(niter) if niter is not None else float("inf"),
            _lib.KX_LS_ACTIVE: 1.0 if (niter is None or niter > 0) else 0.0,
            _lib.KX_LS_VAR: 0.0,
        }
        host = np.zeros(_lib.KX_LS_NSC)
        for k, v in vals.items():
            host[k] = v
        D.sc.copy_(torch.as_tensor(host, device=dev))
        self._cost = torch.zeros((niter if niter is not None else 64) + 1,
                                 dtype=torch.float64, device=dev)
        self._cost[0] = beta
        if show:
            self._print_setup(x)
        if arnorm == 0:
            print(" ")
            print("LSQR finished")
            print(self.msg[self.istop])
        return x

    def _step_body(self, x):
        """Enqueue one LSQR iteration (CB:987-1136) on persistent buffers."""
        D = self._d
        L = D.L
        A = _lib.KX_LS_ACTIVE
        st = D.stream()
        cost = _lib.ptr(self._cost)
        _apply_into(self.Op, self.v, self.tmp_d, False)
        D.vec(_lib.KX_VEC_XPBY, self.tmp_d, self.u, _lib.KX_LS_ALFA, -1.0, _lib.KX_LS_UU, A)
        _lib.check(L.kx_lsqr_scalar(D.s(0), cost, 0, st), "kx_lsqr_scalar")
        D.vec(_lib.KX_VEC_SCALE, None, self.u, _lib.KX_LS_BETA, 1.0, None, A,
              _lib.KX_LS_BETAPOS)
        _apply_into(self.Op, self.u, self.tmp_m, True)
        D.vec(_lib.KX_VEC_XPBY, self.tmp_m, self.v, _lib.KX_LS_BETA, -1.0, _lib.KX_LS_VV, A,
              _lib.KX_LS_BETAPOS)
        _lib.check(L.kx_lsqr_scalar(D.s(0), cost, 1, st), "kx_lsqr_scalar")
        D.vec(_lib.KX_VEC_SCALE, None, self.v, _lib.KX_LS_ALFA, 1.0, None, A,
              _lib.KX_LS_ALFAPOS)
        _lib.check(L.kx_lsqr_scalar(D.s(0), cost, 2, st), "kx_lsqr_scalar")
        rc = L.kx_lsqr_xw(x.numel(), D.kx, _lib.ptr(x), _lib.ptr(self.w), _lib.ptr(self.v),
                          D.s(0), _lib.ptr(D.work), st)
        _lib.check(rc, "kx_lsqr_xw")
        _lib.check(L.kx_lsqr_scalar(D.s(0), cost, 3, st), "kx_lsqr_scalar")

    def step(self, x, show=False):
        """One LSQR iteration (CB:987-1136), enqueued on the device."""
        self._grow(self.iiter + 2)
        self._step_body(x)
        self.iiter += 1
        if show:
            self._print_step(x)
        return x

    def _sync_state(self):
        h = self._d.host()
        self.iiter = int(h[_lib.KX_LS_ITER])
        self.istop = int(h[_lib.KX_LS_ISTOP])
        return h

    def run(self, x, niter=None, show=False, itershow=(10, 10, 10), graph=False):
        """CB:1138-1186.  ``graph=True``: first iteration eager, the next one
        captured in a HIP graph and replayed."""
        niter = self.niter if niter is None else niter
        counts, it0 = _counts(self.Op), self.iiter
        self._grow(niter + 1)
        D = self._d
        D.sc[_lib.KX_LS_NITER] = float(niter)
        D.sc[_lib.KX_LS_ACTIVE] = ((D.sc[_lib.KX_LS_ITER] < niter)
                                   & (D.sc[_lib.KX_LS_ISTOP] == 0)).double()
        if show or self._wants_host_each_step():
            while True:
                h = self._sync_state()
                if not (self.iiter < niter and self.istop == 0):
                    break
                ctol = h[_lib.KX_LS_CTOL]
                showstep = show and (
                    self.niter <= 40 or self.iiter < itershow[0]
                    or niter - self.iiter < itershow[1] or self.iiter % itershow[2] == 0
                    or h[_lib.KX_LS_TEST3] <= 2 * ctol or h[_lib.KX_LS_TEST2] <= 10 * self.atol
                    or h[_lib.KX_LS_TEST1] <= 10 * h[_lib.KX_LS_RTOL])
                x = self.step(x, showstep)
                self.callback(x)
        elif graph and niter - self.iiter > 1:
            x = self.step(x)
            g = _capture(lambda: self._step_body(x), D.device)
            _replay(g, niter - self.iiter, D.sc, _lib.KX_LS_ACTIVE, D.device)
            self._graph = g
            self._sync_state()
        else:
            start = self.iiter
            for k in range(max(niter - start, 0)):
                if k % _CHECK_EVERY == 0 and k > 0:
                    if D.sc[_lib.KX_LS_ACTIVE].item() == 0.0:
                        break
                x = self.step(x)
            self._sync_state()
        _recount(self.Op, counts, self.iiter - it0)
        return x

    def finalize(self, show=False):
        """CB:1188-1201: results copied back from the device scalars."""
        self.tend = time.time()
        self.telapsed = self.tend - self.tstart
        h = self._sync_state()
        self.r1norm = float(h[_lib.KX_LS_R1NORM])
        self.r2norm = float(h[_lib.KX_LS_R2NORM])
        self.anorm = float(h[_lib.KX_LS_ANORM])
        self.acond = float(h[_lib.KX_LS_ACOND])
        self.arnorm = float(h[_lib.KX_LS_ARNORM])
        self.xnorm = float(h[_lib.KX_LS_XNORM])
        # `var` receives the scalar dk.dk on every entry (CB:1066-1069)
        self.var = (torch.full((self.Op.shape[1],), float(h[_lib.KX_LS_VAR]),
                               dtype=torch.float64, device=self._d.device)
                    if self.calc_var else None)
        self.cost = self._cost[:self.iiter + 1].cpu().numpy()
        if show:
            self._print_finalize()

    def solve(self, y, x0=None, damp=0.0, atol=1e-08, btol=1e-08, conlim=100000000.0,
              niter=10, calc_var=True, preallocate=False, show=False, itershow=(10, 10, 10),
              graph=False):
        x = self.setup(y=y, x0=x0, damp=damp, atol=atol, btol=btol, conlim=conlim,
                       niter=niter, calc_var=calc_var, preallocate=preallocate, show=show)
        x = self.run(x, niter=niter, show=show, itershow=itershow, graph=graph)
        self.finalize(show)
        return (x, self.istop, self.iiter, self.r1norm, self.r2norm, self.anorm, self.acond,
                self.arnorm, self.xnorm, self.var, self.cost)
