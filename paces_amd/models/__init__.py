"""Models built on the core frameworks."""

__all__ = ["holstein"]

from . import holstein  # noqa: F401
