from .kirchhoff import Kirchhoff  # noqa: F401
from .kirchhoffbatch import KirchhoffBatch  # noqa: F401
