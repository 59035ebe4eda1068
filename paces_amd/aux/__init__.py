"""Auxiliary functions (paces/aux): row search and helpers."""

from . import helpers, search  # noqa: F401
