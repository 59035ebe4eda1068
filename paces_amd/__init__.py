"""paces_amd: the paces adaptive-basis time step on AMD MI355X (gfx950).

Subpackages mirror the reference package ``paces`` (v1.2.0):
    - config: device selection and verbosity levels
    - core:   Hamiltonian / Observables / TimeEvolution frameworks and the device primitives
    - models: the Holstein model
    - aux:    row search and helpers
The compute runs in libpaces_hip.so (hand-written HIP kernels, C ABI in include/paces_hip.h).
"""

__all__ = ["aux", "core", "models", "config"]
__version__ = "1.2.0+mi355x.1"

from . import config  # noqa: E402,F401
