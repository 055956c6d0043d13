"""Decorators of the operator / solver interface
(pylops/utils/decorators.py:38-211)."""
from functools import wraps

from pylops_amd.config import disabled_ndarray_multiplication

__all__ = ["disable_ndarray_multiplication", "add_ndarray_support_to_solver",
           "reshaped", "count"]


def disable_ndarray_multiplication(func):
    @wraps(func)
    def wrapper(*args, **kwargs):
        with disabled_ndarray_multiplication():
            return func(*args, **kwargs)
    return wrapper


def add_ndarray_support_to_solver(func):
    """Ravel ``y`` (and ``x0``) before the solver runs and reshape the
    solution to ``Op.dims`` unless ``x0`` was given flat or the operator is
    ``forceflat`` (decorators.py:38-70)."""
    @wraps(func)
    def wrapper(A, b, *args, **kwargs):
        x0flat = False
        if "x0" in kwargs and kwargs["x0"] is not None:
            if kwargs["x0"].ndim == 1:
                x0flat = True
            kwargs["x0"] = kwargs["x0"].ravel()
        with disabled_ndarray_multiplication():
            res = list(func(A, b.ravel(), *args, **kwargs))
        if not x0flat and not getattr(A, "forceflat", None):
            res[0] = res[0].reshape(getattr(A, "dims", (A.shape[1],)))
        return tuple(res)
    return wrapper


def reshaped(func=None, forward=None, swapaxis=False):
    """Reshape the input to ``dims`` (forward) or ``dimsd`` (adjoint, any
    function whose name contains "rmat") and ravel the output
    (decorators.py:73-153)."""
    def decorator(f):
        if forward is None:
            fwd = "rmat" not in f.__name__ and f.__name__ not in ("div", "__truediv__")
        else:
            fwd = forward
        inp = "dims" if fwd else "dimsd"

        @wraps(f)
        def wrapper(self, x):
            x = x.reshape(getattr(self, inp))
            if swapaxis:
                x = x.swapaxes(self.axis, -1)
            y = f(self, x)
            if swapaxis:
                y = y.swapaxes(self.axis, -1)
            return y.ravel()
        return wrapper

    if func is not None:
        return decorator(func)
    return decorator


def count(func=None, forward=None, matmat=False):
    """Count forward/adjoint applications (decorators.py:156-211)."""
    def decorator(f):
        fwd = ("rmat" not in f.__name__) if forward is None else forward
        mat = ("matvec" not in f.__name__) if matmat is None else matmat

        @wraps(f)
        def wrapper(self, x):
            y = f(self, x)
            if fwd:
                if mat:
                    self.matmat_count += 1
                    self.matvec_count -= x.shape[-1]
                else:
                    self.matvec_count += 1
            else:
                if mat:
                    self.rmatmat_count += 1
                    self.rmatvec_count -= x.shape[-1]
                else:
                    self.rmatvec_count += 1
            return y
        return wrapper

    if func is not None:
        return decorator(func)
    return decorator
