"""CGLS over a batch of independent problems
(the iteration of pylops/optimization/cls_basic.py:510-564, CGLS.step, run
for every problem of a block-diagonal operator at once).

A plain CGLS on a block-diagonal operator couples its blocks through one
step length per iteration; here every problem keeps its own scalars (step
length, conjugation factor, residual norms, stopping flag), so each one
follows exactly the iterates the reference's CGLS gives it alone.  The
vectors are the concatenations of the problems' vectors; the vector algebra
runs segment-wise (one segment per problem, float64 reductions), the
operator products are the batch's single persistent launches, and the
scalar recurrences are one device thread per problem -- an iteration
enqueues a fixed sequence of kernels with no host synchronisation, which is
captured once in a HIP graph and replayed.
"""
import ctypes
import time

import numpy as np
import torch

from pylops_amd import _lib
from pylops_amd.optimization.cls_basic import _capture

__all__ = ["Segments", "CGLSBatch"]

_CHECK_EVERY = 8


class Segments:
    """A concatenated device vector cut into segments [off[p], off[p+1]),
    and each segment into chunks of at most ``chunk`` elements (one
    workgroup each in the segment-wise kernels)."""

    def __init__(self, offsets, device, chunk=16384):
        off = np.asarray(offsets, dtype=np.int64)
        nseg = len(off) - 1
        lo, seg, first = [], [], []
        for p in range(nseg):
            a, b = int(off[p]), int(off[p + 1])
            first.append(len(lo))
            k = max(1, -(-(b - a) // chunk))
            for j in range(k):
                lo.append(a + j * chunk)
                seg.append(p)
        first.append(len(lo))
        lo.append(int(off[-1]))
        self.nseg, self.nchunk = nseg, len(seg)
        self._off = torch.as_tensor(off, device=device)
        self._chunk_seg = torch.as_tensor(np.asarray(seg, dtype=np.int32), device=device)
        self._chunk_lo = torch.as_tensor(np.asarray(lo, dtype=np.int64), device=device)
        self._seg_chunk = torch.as_tensor(np.asarray(first, dtype=np.int32), device=device)
        self.c = _lib.KxSegs(nseg, self.nchunk, self._off.data_ptr(),
                             self._chunk_seg.data_ptr(), self._chunk_lo.data_ptr(),
                             self._seg_chunk.data_ptr())
        nb = _lib.load().kx_seg_workspace_bytes(ctypes.byref(self.c))
        self.work = torch.zeros((nb + 7) // 8, dtype=torch.float64, device=device)


class CGLSBatch:
    """CGLS for each problem of a batch operator (an operator with
    ``nprob``, ``model_offsets``, ``data_offsets`` and the device product
    ``_into``, such as :class:`KirchhoffBatch`)."""

    def __init__(self, Op):
        self.Op = Op
        self.nprob = Op.nprob

    # ---- launch helpers (on the current stream: a graph capture sees them) ----
    @property
    def _st(self):
        return _lib.stream_ptr(self.device)

    def _vec(self, op, segs, x, y, coef, sign, out, gate=_lib.KX_CG_ACTIVE):
        rc = self._L.kx_seg_vec(op, ctypes.byref(segs.c), self._kx, _lib.ptr(x), _lib.ptr(y),
                                _lib.ptr(self.sc), _lib.KX_CG_NSC, coef, float(sign), out, gate,
                                _lib.ptr(segs.work), self._st)
        _lib.check(rc, "kx_seg_vec")

    def _dot(self, segs, a, b, out, gate=-1):
        rc = self._L.kx_seg_dot(ctypes.byref(segs.c), self._kx, _lib.ptr(a), _lib.ptr(b),
                                _lib.ptr(self.sc), _lib.KX_CG_NSC, out, gate,
                                _lib.ptr(segs.work), self._st)
        _lib.check(rc, "kx_seg_dot")

    def _scalar(self, phase):
        rc = self._L.kx_cgls_scalar_batch(_lib.ptr(self.sc), self.nprob, _lib.ptr(self._cost),
                                          _lib.ptr(self._cost1), self._cost.shape[1], phase,
                                          self._st)
        _lib.check(rc, "kx_cgls_scalar_batch")

    # ---- CB:418-508 per problem ----
    def setup(self, y, x0=None, niter=10, damp=0.0, tol=1e-4):
        """``y``: concatenated data (device tensor); ``damp``: a scalar or one
        value per problem.  Returns the initial concatenated model."""
        Op = self.Op
        self._L = _lib.load()
        dev, dt = y.device, y.dtype
        self.device = dev
        self._kx = _lib.KX_F64 if dt == torch.float64 else _lib.KX_F32
        self.msegs = Segments(Op.model_offsets, dev)
        self.dsegs = Segments(Op.data_offsets, dev)
        self.niter = int(niter)
        self.tol = float(tol)
        damp = np.broadcast_to(np.asarray(damp, dtype=np.float64), (self.nprob,)).copy()
        self.damp = damp
        n = self.nprob
        self.sc = torch.zeros((n, _lib.KX_CG_NSC), dtype=torch.float64, device=dev)
        y = y.reshape(-1).contiguous()
        self.y = y
        nm = Op.shape[1]
        mlen = np.diff(Op.model_offsets)
        dampm = torch.as_tensor(np.repeat(damp, mlen), dtype=dt, device=dev)
        self.c = torch.empty(nm, dtype=dt, device=dev)
        if x0 is None:
            x = torch.zeros(nm, dtype=dt, device=dev)
            self.s = y.clone()
            Op._into(self.s, self.c, True)
        else:
            x = x0.reshape(-1).to(device=dev, dtype=dt).clone()
            q0 = torch.empty(Op.shape[0], dtype=dt, device=dev)
            Op._into(x, q0, False)
            self.s = y - q0
            Op._into(self.s, self.c, True)
            self.c -= dampm * x  # raw damp, as CB:477
        self.q = torch.empty(Op.shape[0], dtype=dt, device=dev)
        self.r = torch.empty(nm, dtype=dt, device=dev)
        Op._into(self.c, self.q, False)
        self._dot(self.msegs, self.c, self.c, _lib.KX_CG_KOLD)
        self._dot(self.msegs, self.c, self.c, _lib.KX_CG_CC)
        self._dot(self.dsegs, self.q, self.q, _lib.KX_CG_QQ)
        self._dot(self.dsegs, self.s, self.s, _lib.KX_CG_SS)
        self._dot(self.msegs, x, x, _lib.KX_CG_XX)
        sc = self.sc
        self._cost = torch.zeros((n, self.niter + 1), dtype=torch.float64, device=dev)
        self._cost1 = torch.zeros((n, self.niter + 1), dtype=torch.float64, device=dev)
        dampd = torch.as_tensor(damp, device=dev)
        self._cost[:, 0] = torch.sqrt(sc[:, _lib.KX_CG_SS])
        self._cost1[:, 0] = torch.sqrt(sc[:, _lib.KX_CG_SS] + dampd * sc[:, _lib.KX_CG_XX].abs())
        sc[:, _lib.KX_CG_DAMP2] = dampd ** 2
        sc[:, _lib.KX_CG_TOL] = self.tol
        sc[:, _lib.KX_CG_ITER] = 0.0
        sc[:, _lib.KX_CG_NITER] = float(self.niter)
        sc[:, _lib.KX_CG_ACTIVE] = (sc[:, _lib.KX_CG_KOLD] > self.tol).double()
        self._any_damp = bool(np.any(damp != 0))
        self.iiter = 0
        return x

    # ---- CB:510-564 per problem ----
    def _step_body(self, x):
        """Enqueue one iteration for every problem; touches only persistent
        buffers, so it can be captured and replayed.  A problem whose loop
        flag is off is left untouched."""
        Op, ms, ds = self.Op, self.msegs, self.dsegs
        self._scalar(0)
        self._vec(_lib.KX_VEC_AXPY, ms, self.c, x, _lib.KX_CG_A, 1.0, _lib.KX_CG_XX)
        self._vec(_lib.KX_VEC_AXPY, ds, self.q, self.s, _lib.KX_CG_A, -1.0, _lib.KX_CG_SS)
        Op._into(self.s, self.r, True)
        if self._any_damp:
            self._vec(_lib.KX_VEC_AXPY, ms, x, self.r, _lib.KX_CG_DAMP2, -1.0, _lib.KX_CG_K)
        else:
            self._dot(ms, self.r, self.r, _lib.KX_CG_K, _lib.KX_CG_ACTIVE)
        self._scalar(1)
        self._vec(_lib.KX_VEC_XPBY, ms, self.r, self.c, _lib.KX_CG_B, 1.0, _lib.KX_CG_CC)
        Op._into(self.c, self.q, False)
        self._dot(ds, self.q, self.q, _lib.KX_CG_QQ, _lib.KX_CG_ACTIVE)
        self._scalar(2)

    def _any_active(self):
        return bool((self.sc[:, _lib.KX_CG_ACTIVE] != 0).any().item())

    def run(self, x, graph=True):
        """Iterate up to niter times per problem.  ``graph``: the first
        iteration runs eagerly, the next one is captured in a HIP graph and
        replayed for the rest."""
        todo = self.niter
        if graph and todo > 1:
            self._step_body(x)
            g = _capture(lambda: self._step_body(x), self.device)
            for k in range(todo - 1):
                if k % _CHECK_EVERY == 0 and k > 0 and not self._any_active():
                    break
                g.replay()
            self._graph = g
        else:
            for k in range(todo):
                if k % _CHECK_EVERY == 0 and k > 0 and not self._any_active():
                    break
                self._step_body(x)
        return x

    def finalize(self):
        """Per problem, as CB:613-630: istop from kold vs tol, r1norm = kold,
        r2norm = cost1[iiter], cost = cost[:iiter + 1]."""
        h = self.sc.cpu().numpy()
        cost = self._cost.cpu().numpy()
        cost1 = self._cost1.cpu().numpy()
        self.iiter = h[:, _lib.KX_CG_ITER].astype(np.int64)
        self.kold = np.abs(h[:, _lib.KX_CG_KOLD])
        self.istop = np.where(self.kold < self.tol, 1, 2)
        self.r1norm = self.kold
        self.r2norm = cost1[np.arange(self.nprob), self.iiter]
        self.cost = [cost[p, :self.iiter[p] + 1] for p in range(self.nprob)]
        self.cost1 = [cost1[p, :self.iiter[p] + 1] for p in range(self.nprob)]

    def solve(self, y, x0=None, niter=10, damp=0.0, tol=1e-4, graph=True):
        self.tstart = time.time()
        x = self.setup(y, x0=x0, niter=niter, damp=damp, tol=tol)
        x = self.run(x, graph=graph)
        self.finalize()
        self.tend = time.time()
        self.telapsed = self.tend - self.tstart
        return x
