"""s2fft_amd: MI355X-native spin spherical-harmonic and Wigner transforms.

Drop-in ``method="hip"`` implementations of s2fft.forward / s2fft.inverse
(s2fft/__init__.py:5-19) backed by hand-written HIP kernels for gfx950.
"""

from . import precompute_transforms, sampling  # noqa: F401
from .transforms import spherical, wigner  # noqa: F401
from .transforms.spherical import forward, inverse  # noqa: F401

__all__ = ["forward", "inverse", "spherical", "wigner", "sampling", "precompute_transforms"]
