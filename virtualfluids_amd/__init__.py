"""virtualfluids_amd — MI355X-native (gfx950) K17 lattice-Boltzmann time step.

The compute path is hand-written HIP in csrc/, exported through the C ABI of
include/virtualfluids_amd.h (libvfamd.so). This package is the Python side of that
boundary: grid builders, device state, the time-step driver and the autograd bridge.
"""
from . import capi, grid  # noqa: F401

__all__ = ["capi", "grid"]
