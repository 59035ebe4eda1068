"""Global configuration (paces/config/__init__.py)."""

from .dbg_verbosity_levels import *  # noqa: F401,F403
from ._dev_conf import DeviceConfig

devices = DeviceConfig()
