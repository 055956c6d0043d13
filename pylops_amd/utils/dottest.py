"""Adjointness check (pylops/utils/dottest.py:10-132), real operators."""
import numpy as np

__all__ = ["dottest"]


def dottest(Op, nr=None, nc=None, rtol=1e-6, atol=1e-21, complexflag=0,
            raiseerror=True, verb=False, backend="numpy"):
    """(Op u)^H v == u^H (Op^H v) within rtol/atol on N(0,1) vectors of the
    operator's real dtype.  ``backend="torch"`` draws the vectors as device
    tensors on the current HIP device."""
    if nr is None:
        nr = Op.shape[0]
    if nc is None:
        nc = Op.shape[1]
    if (nr, nc) != Op.shape:
        raise AssertionError("Provided nr and nc do not match operator shape")
    if complexflag != 0:
        raise NotImplementedError("complex dottest is not supported")
    rdtype = np.ones(1, Op.dtype).real.dtype
    u = np.random.randn(nc).astype(rdtype)
    v = np.random.randn(nr).astype(rdtype)
    if backend == "torch":
        import torch
        u = torch.as_tensor(u, device="cuda")
        v = torch.as_tensor(v, device="cuda")
    y = Op.matvec(u)
    x = Op.rmatvec(v)
    if backend == "torch":
        yy = float((y.double() * v.double()).sum())
        xx = float((u.double() * x.double()).sum())
    else:
        yy = np.vdot(np.asarray(y), v)
        xx = np.vdot(u, np.asarray(x))
    passed = np.isclose(xx, yy, rtol, atol)
    if (not passed and raiseerror) or verb:
        status = "passed" if passed else "failed"
        msg = f"Dot test {status}, v^H(Opu)={yy} - u^H(Op^Hv)={xx}"
        if not passed and raiseerror:
            raise AssertionError(msg)
        print(msg)
    return passed
