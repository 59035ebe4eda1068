"""Auxiliary functions (paces/aux): row search, helpers and the scaling-and-squaring
propagator (paces/aux/__init__.py:8-11)."""

from . import helpers, search  # noqa: F401
from .expm import expm_multiply, expm_multiply_simple  # noqa: F401
