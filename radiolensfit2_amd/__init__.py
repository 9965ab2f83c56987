"""MI355X-native per-galaxy shape measurement for radio weak lensing.

A from-scratch gfx950 implementation of RadioLensfit2's per-galaxy path (source extraction,
exponential-disk likelihood with flux / position / scalelength marginalisation, simplex fit
and likelihood sampling) behind a C ABI that keeps the reference's function surface.
Native code: librl2hip.so (HIP kernels + runtime), librl2sim.so (synthetic observations).
"""
from . import _native  # noqa: F401

__all__ = ["_native", "synth", "pipeline", "build"]
