"""Eikonal solver front end: dictIn / dictOut (agd/Eikonal/__init__.py)."""
from .DictIn import dictIn, dictOut, CenteredLinspace  # noqa: F401
from . import autograd  # noqa: F401,E402
