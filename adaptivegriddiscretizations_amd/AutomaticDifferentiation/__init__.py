"""First-order forward AD carrier of the eikonal front end.

Only what the solver's forward-AD path exchanges with its caller
(agd/AutomaticDifferentiation/Dense.py:14-44 denseAD, and is_ad /
remove_ad of agd/AutomaticDifferentiation/__init__.py): a value array plus
a trailing axis of first-order coefficients.  Arithmetic on AD variables is
not provided.
"""
from . import Dense  # noqa: F401
from .Dense import denseAD  # noqa: F401


def is_ad(x):
    """True for an AD variable (a denseAD)."""
    return isinstance(x, Dense.denseAD)


def remove_ad(x):
    """The value of an AD variable, x itself otherwise."""
    return x.value if is_ad(x) else x
