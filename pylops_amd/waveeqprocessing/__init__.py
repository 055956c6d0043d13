from .kirchhoff import Kirchhoff  # noqa: F401
from .kirchhoffbatch import KirchhoffBatch  # noqa: F401
from .autotune import autotune, autotune_batch  # noqa: F401
from .lsm import LSM  # noqa: F401
