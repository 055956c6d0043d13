from .convolve1d import Convolve1D  # noqa: F401
