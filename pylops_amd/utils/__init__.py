from .dottest import dottest  # noqa: F401
from .wavelets import ricker  # noqa: F401
from .tapers import taper, hanningtaper  # noqa: F401
