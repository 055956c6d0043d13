"""Many independent Kirchhoff operators applied as one.

``KirchhoffBatch(ops)`` is the block-diagonal operator of a list of
:class:`Kirchhoff` operators: its products equal those of the reference's
``BlockDiag(ops)`` (pylops/basicoperators/blockdiag.py:182-245) on the
concatenated models and data.  Where BlockDiag applies the operators one
after the other, every product here is one persistent launch per engine
kernel over all problems (``kx_batch_forward`` / ``kx_batch_adjoint`` of
libkirchx): work items of all problems, largest first, are pulled from one
device counter, so problems of uneven sizes share the chip without a launch
per problem.  Each problem's result is that of its own operator's kernels.
"""
import ctypes

import numpy as np
import torch

from pylops_amd import _lib
from pylops_amd.linearoperator import LinearOperator
from pylops_amd.utils.backend import like_input, to_device
from pylops_amd.waveeqprocessing.kirchhoff import Kirchhoff

__all__ = ["KirchhoffBatch"]


class KirchhoffBatch(LinearOperator):
    """Block-diagonal batch of Kirchhoff operators on one device.

    Parameters
    ----------
    ops : list of :class:`Kirchhoff`
        Operators of one dtype, table format, ``dynamic`` flag and device.
    rblock : int
        Receivers per forward workgroup for the whole batch (0 = automatic).
    sorder : int
        Forward lane order, ``_lib.KX_ORDER_*`` (0 = automatic).
    gsplit : int
        Adjoint waves per tile group, 1, 2 or 4 (0 = automatic).
    gtiles : int
        Adjoint tiles per wave, 1 or 2 (0 = automatic).
    """

    def __init__(self, ops, rblock=0, sorder=0, gsplit=0, gtiles=0, name="B"):
        ops = list(ops)
        if not ops:
            raise ValueError("KirchhoffBatch needs at least one operator")
        for op in ops:
            if not isinstance(op, Kirchhoff):
                raise TypeError("KirchhoffBatch takes Kirchhoff operators")
        dt = ops[0].dtype
        dev = ops[0].device
        for op in ops:
            if op.dtype != dt:
                raise ValueError("every operator of a batch must have the same dtype")
            if op.device != dev:
                raise ValueError("every operator of a batch must live on the same device")
        self.ops = ops
        self.nprob = len(ops)
        self.device = dev
        n = self.nprob
        plans = (_lib.KxPlan * n)()
        for i, op in enumerate(ops):
            plans[i] = op.plan
        plans[0].rblock = int(rblock)
        plans[0].sorder = int(sorder)
        plans[0].gsplit = int(gsplit)
        plans[0].gtiles = int(gtiles)
        L = _lib.load()
        nbytes = L.kx_batch_dev_bytes(plans, n)
        if nbytes < 0:
            _lib.check(int(nbytes), "kx_batch_dev_bytes")
        # the device tables of every operator stay alive with the batch, even
        # if an operator rebuilds its own
        self._keep = [op._dev_tables for op in ops]
        self._dev = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
        self._batch = _lib.KxBatch()
        with torch.cuda.device(dev):
            rc = L.kx_batch_build(ctypes.byref(self._batch), plans, n, _lib.ptr(self._dev),
                                  _lib.stream_ptr(dev))
        _lib.check(rc, "kx_batch_build")
        self.model_offsets = np.concatenate(([0], np.cumsum([op.ni for op in ops])))
        self.data_offsets = np.concatenate(([0], np.cumsum([op.shape[0] for op in ops])))
        assert self.model_offsets[-1] == self._batch.model_len
        assert self.data_offsets[-1] == self._batch.data_len
        nd, nm = int(self.data_offsets[-1]), int(self.model_offsets[-1])
        super().__init__(dtype=dt, shape=(nd, nm), dims=(nm,), dimsd=(nd,), name=name)
        tdt = torch.float64 if dt == np.float64 else torch.float32
        self._ws = torch.empty(nd, dtype=tdt, device=dev)

    @property
    def rblock(self):
        return self._batch.rblock

    @property
    def gsplit(self):
        return self._batch.gsplit

    @property
    def gtiles(self):
        return self._batch.gtiles

    @property
    def sorder(self):
        return self._batch.sorder

    def split_model(self, m):
        """Views of the concatenated model, one per problem."""
        o = self.model_offsets
        return [m[o[i]:o[i + 1]] for i in range(self.nprob)]

    def split_data(self, d):
        """Views of the concatenated data, one per problem."""
        o = self.data_offsets
        return [d[o[i]:o[i + 1]] for i in range(self.nprob)]

    def _into(self, x, out, adjoint):
        """out <- B x (B^H x if adjoint) for contiguous device tensors; the
        solvers' fast path (no allocation, so it can be graph-captured)."""
        L = _lib.load()
        fn = L.kx_batch_adjoint if adjoint else L.kx_batch_forward
        rc = fn(ctypes.byref(self._batch), _lib.ptr(x), _lib.ptr(out), _lib.ptr(self._ws),
                _lib.stream_ptr(x.device))
        _lib.check(rc, "kx_batch_adjoint" if adjoint else "kx_batch_forward")
        if adjoint:
            self.rmatvec_count += 1
        else:
            self.matvec_count += 1
        return out

    def _launch(self, x, nout, adjoint):
        xd = to_device(x.reshape(-1), self.dtype, self.device).contiguous()
        out = torch.empty(nout, dtype=xd.dtype, device=self.device)
        L = _lib.load()
        fn = L.kx_batch_adjoint if adjoint else L.kx_batch_forward
        rc = fn(ctypes.byref(self._batch), _lib.ptr(xd), _lib.ptr(out), _lib.ptr(self._ws),
                _lib.stream_ptr(self.device))
        _lib.check(rc, "kx_batch_adjoint" if adjoint else "kx_batch_forward")
        return like_input(out, x)

    def _matvec(self, x):
        return self._launch(x, self.shape[0], False)

    def _rmatvec(self, x):
        return self._launch(x, self.shape[1], True)
