"""pylops_amd: the PyLops Kirchhoff + CGLS/LSQR path on AMD MI355X.

Drop-in names: ``LinearOperator``, ``waveeqprocessing.Kirchhoff``,
``signalprocessing.Convolve1D``, ``optimization.basic.cgls / lsqr``,
``utils.dottest``, ``TorchOperator``.  Compute runs in the HIP library
libkirchx.so (built by ``python -m pylops_amd.build``).
"""
from .config import (  # noqa: F401
    disabled_ndarray_multiplication, enabled_ndarray_multiplication,
    get_ndarray_multiplication, set_ndarray_multiplication)
from .linearoperator import LinearOperator, aslinearoperator  # noqa: F401
from . import utils  # noqa: F401
from . import signalprocessing  # noqa: F401
from . import waveeqprocessing  # noqa: F401
from . import optimization  # noqa: F401
from .signalprocessing import Convolve1D  # noqa: F401
from .waveeqprocessing import Kirchhoff  # noqa: F401
from .torchoperator import TorchOperator  # noqa: F401

__version__ = "0.1.0"
