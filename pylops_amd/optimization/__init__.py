from . import basic  # noqa: F401
from .basic import cgls, cgls_batch, lsqr  # noqa: F401
from .callback import Callbacks  # noqa: F401
from .cls_basic import CGLS, LSQR  # noqa: F401
from .cls_batch import CGLSBatch  # noqa: F401
from . import sparsity  # noqa: F401
from .cls_sparsity import FISTA, ISTA  # noqa: F401
from .sparsity import fista, ista  # noqa: F401
