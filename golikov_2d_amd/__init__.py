"""MI355X-native explicit finite-volume step for 2-D Euler / ideal MHD.

The compute path is libgd2.so (hand-written HIP kernels for gfx950 behind a
C ABI, declared in include/gd2.h); this package holds the host-side Python
mirror of that interface and the synthetic problem generators.
"""
__all__ = ["problems"]
