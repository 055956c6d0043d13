from .kirchhoff import Kirchhoff  # noqa: F401
