"""Function front-ends of the solvers (pylops/optimization/basic.py:89-332):
same signatures and return tuples as the reference.  numpy inputs are moved
to the operator's device and results come back as numpy; torch device
tensors stay on the device."""
import numpy as np
import torch

from pylops_amd.optimization.cls_basic import CGLS, LSQR
from pylops_amd.utils.backend import is_torch, torch_dtype
from pylops_amd.utils.decorators import add_ndarray_support_to_solver

__all__ = ["cgls", "lsqr", "cgls_batch"]


def _device_of(Op, y):
    if is_torch(y) and y.is_cuda:
        return y.device
    dev = getattr(Op, "device", None)
    return dev if dev is not None else torch.device("cuda", torch.cuda.current_device())


def _to_dev(a, device, dtype, flat=True):
    if a is None:
        return None
    if is_torch(a):
        a = a.to(device=device, dtype=dtype)
        return (a.reshape(-1) if flat else a).contiguous()
    a = np.asarray(a)
    return torch.as_tensor(np.ascontiguousarray(a.ravel() if flat else a), dtype=dtype,
                           device=device)


def _prepare(Op, y, x0, callback, solver, flat=True):
    """y and x0 on the operator's device (raveled unless ``flat`` is False),
    the callback wrapped for numpy callers, and the map back to the
    caller's array type."""
    from pylops_amd import _lib
    _lib.require_device()
    dev = _device_of(Op, y)
    dt = y.dtype if is_torch(y) else torch_dtype(np.asarray(y).dtype)
    numpy_io = not is_torch(y)
    yd = _to_dev(y, dev, dt, flat)
    x0d = _to_dev(x0, dev, dt, flat)
    if callback is not None:
        if numpy_io:
            solver.callback = lambda x: callback(x.detach().cpu().numpy())
        else:
            solver.callback = callback
    back = (lambda t: t.detach().cpu().numpy()) if numpy_io else (lambda t: t)
    return yd, x0d, back


@add_ndarray_support_to_solver
def cgls(Op, y, x0=None, niter=10, damp=0.0, tol=1e-4, show=False, itershow=(10, 10, 10),
         callback=None, preallocate=False, graph=False):
    """Conjugate gradient least squares; returns
    ``(x, istop, iit, r1norm, r2norm, cost)`` as the reference.  ``graph``
    replays a HIP-graph capture of one iteration."""
    cgsolve = CGLS(Op)
    yd, x0d, back = _prepare(Op, y, x0, callback, cgsolve)
    x, istop, iiter, r1norm, r2norm, cost = cgsolve.solve(
        y=yd, x0=x0d, tol=tol, niter=niter, damp=damp, show=show, itershow=itershow,
        preallocate=preallocate, graph=graph)
    return back(x), istop, iiter, r1norm, r2norm, cost


@add_ndarray_support_to_solver
def lsqr(Op, y, x0=None, damp=0.0, atol=1e-08, btol=1e-08, conlim=100000000.0, niter=10,
         calc_var=True, show=False, itershow=(10, 10, 10), callback=None, preallocate=False,
         graph=False):
    """LSQR; returns ``(x, istop, iiter, r1norm, r2norm, anorm, acond,
    arnorm, xnorm, var, cost)`` as the reference.  ``graph`` replays a
    HIP-graph capture of one iteration."""
    lsqrsolve = LSQR(Op)
    yd, x0d, back = _prepare(Op, y, x0, callback, lsqrsolve)
    (x, istop, iiter, r1norm, r2norm, anorm, acond, arnorm, xnorm, var,
     cost) = lsqrsolve.solve(y=yd, x0=x0d, damp=damp, atol=atol, btol=btol, conlim=conlim,
                             niter=niter, calc_var=calc_var, show=show, itershow=itershow,
                             preallocate=preallocate, graph=graph)
    if var is not None:
        var = back(var)
    return back(x), istop, iiter, r1norm, r2norm, anorm, acond, arnorm, xnorm, var, cost


def cgls_batch(ops, ys, x0s=None, niter=10, damp=0.0, tol=1e-4, graph=True):
    """CGLS on many independent problems at once: ``ops`` is a list of
    Kirchhoff operators (or a ``KirchhoffBatch``), ``ys`` their data (a list,
    or the concatenated data), ``x0s`` optional starting models likewise;
    ``damp`` a scalar or one value per problem.  Returns one tuple
    ``(x, istop, iit, r1norm, r2norm, cost)`` per problem, each as the
    reference's ``cgls`` (basic.py:89-178) returns for that problem alone.
    Models come back as numpy arrays when the data were numpy arrays."""
    from pylops_amd import _lib
    from pylops_amd.optimization.cls_batch import CGLSBatch
    from pylops_amd.waveeqprocessing.kirchhoffbatch import KirchhoffBatch
    _lib.require_device()
    B = ops if isinstance(ops, KirchhoffBatch) else KirchhoffBatch(ops)
    dev = B.device
    dt = torch.float64 if B.dtype == np.float64 else torch.float32

    def cat(vs):
        if isinstance(vs, (list, tuple)):
            numpy_io = not all(is_torch(v) for v in vs)
            return torch.cat([_to_dev(v, dev, dt) for v in vs]), numpy_io
        return _to_dev(vs, dev, dt), not is_torch(vs)

    yd, numpy_io = cat(ys)
    if yd.numel() != B.shape[0]:
        raise ValueError(f"data of {yd.numel()} elements for a batch of shape {B.shape}")
    x0d = None if x0s is None else cat(x0s)[0]
    solver = CGLSBatch(B)
    x = solver.solve(yd, x0=x0d, niter=niter, damp=damp, tol=tol, graph=graph)
    out = []
    for p, xp in enumerate(B.split_model(x)):
        xp = xp.detach().cpu().numpy() if numpy_io else xp
        out.append((xp, int(solver.istop[p]), int(solver.iiter[p]), float(solver.r1norm[p]),
                    float(solver.r2norm[p]), solver.cost[p]))
    return out
