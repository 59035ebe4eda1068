"""paces: the reference's import path (paces/__init__.py:1-16) for the MI355X engine.

Scripts written for the reference import ``paces.config``, ``paces.models.holstein`` and
``paces.core.time_evolution`` (main_example.py:8-10). This package makes each of those names the
very module object of ``paces_amd`` that implements it, so classes, the device configuration and
the loaded HIP library are shared whichever name a caller uses:

    from paces.config import devices                  # paces_amd.config
    from paces.models import holstein                 # paces_amd.models.holstein
    from paces.core.time_evolution import ExpmParams  # paces_amd.core.time_evolution
    from paces.aux import expm_multiply_simple        # paces_amd.aux.expm_multiply_simple
"""

import importlib
import sys

import paces_amd as _impl

__all__ = ["aux", "core", "models", "config"]
__version__ = _impl.__version__

# reference module name -> implementing module (the reference's cupy_search is aux.search here)
_MODULES = {
    "config": "paces_amd.config",
    "config._dev_conf": "paces_amd.config._dev_conf",
    "config.dbg_verbosity_levels": "paces_amd.config.dbg_verbosity_levels",
    "core": "paces_amd.core",
    "core.hamiltonian": "paces_amd.core.hamiltonian",
    "core.observables": "paces_amd.core.observables",
    "core.time_evolution": "paces_amd.core.time_evolution",
    "models": "paces_amd.models",
    "models.holstein": "paces_amd.models.holstein",
    "aux": "paces_amd.aux",
    "aux.helpers": "paces_amd.aux.helpers",
    "aux.cupy_search": "paces_amd.aux.search",
    "aux.expm": "paces_amd.aux.expm",
}

for _name, _target in _MODULES.items():
    _mod = importlib.import_module(_target)
    sys.modules[f"{__name__}.{_name}"] = _mod
    _parent, _, _leaf = _name.rpartition(".")
    setattr(sys.modules[f"{__name__}.{_parent}"] if _parent else sys.modules[__name__], _leaf,
            _mod)

del _name, _target, _mod, _parent, _leaf
