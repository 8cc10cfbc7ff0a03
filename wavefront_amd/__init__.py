"""wavefront_amd: MI355X-native timestep engine for the wavefront FDTD solver.

The device path (libwavefront_hip.so, hand-written HIP for gfx950) advances
the Yee H-then-E updates with fused graded CPML, soft/hard point sources and
running-DFT probes; this package mirrors the reference's Rust API
(layout / source / probe / scene / engine) on top of its C ABI.
"""
from ._lib import WavefrontError, lib  # noqa: F401

__all__ = ["WavefrontError", "lib"]
