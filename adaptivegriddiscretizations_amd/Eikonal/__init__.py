"""Eikonal solver front end: dictIn / dictOut (agd/Eikonal/__init__.py)."""
from .DictIn import dictIn, dictOut, CenteredLinspace  # noqa: F401
