"""Module-level behaviour flags (pylops/config.py): ndarray multiplication."""
from contextlib import contextmanager

__all__ = [
    "get_ndarray_multiplication",
    "set_ndarray_multiplication",
    "enabled_ndarray_multiplication",
    "disabled_ndarray_multiplication",
]

_ndarray_multiplication = True


def get_ndarray_multiplication() -> bool:
    return _ndarray_multiplication


def set_ndarray_multiplication(val: bool) -> None:
    global _ndarray_multiplication
    _ndarray_multiplication = bool(val)


@contextmanager
def enabled_ndarray_multiplication():
    old = get_ndarray_multiplication()
    set_ndarray_multiplication(True)
    try:
        yield old
    finally:
        set_ndarray_multiplication(old)


@contextmanager
def disabled_ndarray_multiplication():
    old = get_ndarray_multiplication()
    set_ndarray_multiplication(False)
    try:
        yield old
    finally:
        set_ndarray_multiplication(old)
