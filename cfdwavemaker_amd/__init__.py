"""MI355X-native second-order irregular-wave kinematics (CFDwavemaker drop-in).

The compute path is libcfdwm.so (C++ host code + hand-written gfx950 HIP
kernels) behind the reference's C ABI; this package holds its Python
bindings and the sea-state generator.
"""
from .seastate import SeaState, make_sea, config_sea, write_waveinput, wavenumber  # noqa: F401


def lib():
    """Load libcfdwm.so (raises if it has not been built)."""
    from ._lib import load
    return load()
