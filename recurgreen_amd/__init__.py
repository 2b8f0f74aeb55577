"""recurgreen_amd -- MI355X (gfx950) recursive Green function sweep of RecurGreen's waveguide path.

The compute runs in librecurgreen_hip.so (HIP kernels, C ABI in include/recurgreen_hip.h);
this package holds the host-side mirror of the reference interface.
"""
from .binding import Plan, RecurGreenError, lib  # noqa: F401
from .waveguide import Waveguide  # noqa: F401

__all__ = ["Plan", "RecurGreenError", "Waveguide", "lib"]
